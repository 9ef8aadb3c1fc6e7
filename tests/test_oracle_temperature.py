"""Element temperatures in the oracle: OpenMC's "nearest" table per element.

The oracle looks the table up per segment from the element's sqrt(kT)
(CollisionKernel.C:109, OpenMC Nuclide::calculate_xs).  These tests pin that
against problems with no temperatures at all that must give the same
histories: every element's choice is worked out here with numpy, then each
(nuclide, table) becomes a nuclide of its own and each element's (material,
tables) a material of its own, so nothing temperature-related is left for
the oracle to do."""
import dataclasses

import numpy as np
import pytest

from macaw_amd import abi, problems as P
from oracle.pyoracle import Oracle, OracleError


def small_c2(n_lattice=8, n_particles=1500, n_points=3000):
    return P.c2_tet_vacuum(n_lattice=n_lattice, n_particles=n_particles, n_points=n_points,
                           tallies=nuclide_tallies())


def nuclide_tallies():
    return [P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                        scores=[abi.SCORE_FLUX, abi.SCORE_TOTAL, abi.SCORE_ABSORPTION]),
            P.TallySpec(id=2, estimator=abi.EST_COLLISION, filters=[], scores=[abi.SCORE_TOTAL],
                        nuclides=[-1])]


def tables_of(p, q):
    """Nuclide q's tables in ascending temperature."""
    ts = [p.nuclides[q]] + list(p.nuclide_temps.get(q, []))
    temp = lambda t: t.temperature if t.temperature > 0 else 293.6
    return sorted(ts, key=temp), [temp(t) for t in sorted(ts, key=temp)]


def nearest(temps, T):
    """OpenMC 'nearest': argmin over ascending tables of |k T_t - sqrt(k T)^2|,
    first minimum (the lower temperature) on a tie."""
    skT = np.sqrt(abi.K_BOLTZMANN * T)
    kT = skT * skT
    d = np.abs(abi.K_BOLTZMANN * np.asarray(temps)[None, :] - kT[:, None])
    return np.argmin(d, axis=1)


def expand(p):
    """The same problem without temperatures: one nuclide per (nuclide,
    table) in use, one material (and block) per (material, choices)."""
    mesh = p.mesh
    id_to_mat = {m.id: m for m in p.materials}
    emat = [id_to_mat[p.block_material[int(b)]] for b in mesh.subdomain]
    choice = {}
    for q in range(len(p.nuclides)):
        _, temps = tables_of(p, q)
        choice[q] = nearest(temps, p.elem_temperature)
    nucs, nuc_id, mats, mat_id = [], {}, [], {}
    sub = np.zeros(mesh.n_elem, np.int32)
    for e in range(mesh.n_elem):
        m = emat[e]
        ids = []
        for q in m.nuclides:
            key = (int(q), int(choice[int(q)][e]))
            if key not in nuc_id:
                nuc_id[key] = len(nucs)
                nucs.append(tables_of(p, key[0])[0][key[1]])
            ids.append(nuc_id[key])
        mk = (m.id, tuple(ids))
        if mk not in mat_id:
            mat_id[mk] = len(mats)
            mats.append(P.MaterialData(id=len(mats) + 1, nuclides=np.array(ids, np.int32),
                                       densities=np.asarray(m.densities, np.float64)))
        sub[e] = mat_id[mk]
    mesh2 = dataclasses.replace(mesh, subdomain=sub)
    return dataclasses.replace(p, name=p.name + "_expanded", mesh=mesh2, nuclides=nucs, materials=mats,
                               block_material={b: b + 1 for b in range(len(mats))}, nuclide_temps={},
                               elem_temperature=None), len(nucs), len(mats)


def run(p, n=None, batches=1, **kw):
    n = n or p.n_particles
    o = Oracle(p, **kw)
    for b in range(1, batches + 1):
        o.run_batch(b, n, threads=1, record=True)
    return o, o.histories(n)


def assert_same_histories(a, b):
    for f in a.dtype.names:
        x, y = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        assert x.tobytes() == y.tobytes(), f


def test_nearest_table_equals_expanded_problem():
    p = P.with_temperatures(small_c2())
    q, n_nuc, n_mat = expand(p)
    # the field spans every table and several choices per material
    assert n_nuc == 4 * len(P.TEMPERATURES) and n_mat > 4
    oa, ha = run(p)
    ob, hb = run(q)
    assert_same_histories(ha, hb)
    # flux and total-material scores per element are the same sums
    ta, tb = oa.tally(1).reshape(-1, 3), ob.tally(1).reshape(-1, 3)
    np.testing.assert_array_equal(ta[:, 0], tb[:, 0])
    np.testing.assert_allclose(ta, tb, rtol=1e-13, atol=0)
    assert ta[:, 1].sum() > 0
    np.testing.assert_allclose(oa.tally(2), ob.tally(2), rtol=1e-12, atol=0)
    # and the temperature changes the answer
    oc, hc = run(dataclasses.replace(p, elem_temperature=None))
    assert ha["digest"].tobytes() != hc["digest"].tobytes()


def test_eigenvalue_with_temperatures_equals_expanded_problem():
    p = P.with_temperatures(P.k2_tet_eigen(n_lattice=7, n_particles=1200, n_points=3000), axis=2)
    q, _, _ = expand(p)
    oa, ha = run(p, batches=2, eigenvalue=True, n_inactive=1)
    ob, hb = run(q, batches=2, eigenvalue=True, n_inactive=1)
    assert_same_histories(ha, hb)
    ka, kb = oa.keff(), ob.keff()
    for k in ("fission_sites", "k_generation", "k_collision", "k_absorption", "keff"):
        assert ka[k] == kb[k], k
    assert oa.source_bank().tobytes() == ob.source_bank().tobytes()


def test_uniform_temperature_is_that_table():
    p = P.with_temperatures(small_c2(n_particles=800))
    hot = dataclasses.replace(p, elem_temperature=np.full(p.mesh.n_elem, 905.0))
    only = dataclasses.replace(p, nuclides=[ts[1] for ts in (p.nuclide_temps[q] for q in range(4))],
                               nuclide_temps={}, elem_temperature=None)
    assert all(t.temperature == 900.0 for t in only.nuclides)
    assert_same_histories(run(hot)[1], run(only)[1])


def test_without_element_temperatures_first_table_is_read():
    # the first table need not be the coldest: it is read everywhere
    p = small_c2(n_particles=800)
    hot = P.c2_nuclides(3000, 4, temperature=1200.0)
    cold = P.c2_nuclides(3037, 4, temperature=293.6)
    a = dataclasses.replace(p, nuclides=hot, nuclide_temps={q: [cold[q]] for q in range(4)})
    b = dataclasses.replace(p, nuclides=hot)
    assert_same_histories(run(a)[1], run(b)[1])


def test_nearest_rule_boundaries():
    temps = [293.6, 600.0, 900.0]
    T = np.array([0.0, 293.6, 446.0, 447.0, 600.0, 749.9, 750.1, 5000.0])
    # |600 - 446| = 154 > |293.6 - 446| = 152.4 -> 293.6; 447 -> 600
    assert nearest(temps, T).tolist() == [0, 0, 0, 1, 1, 1, 2, 2]


def test_temperature_out_of_tolerance_is_an_error():
    p = P.with_temperatures(small_c2(n_particles=100), tolerance=10.0)
    o = Oracle(p)
    with pytest.raises(OracleError, match="no cross sections"):
        o.run_batch(1, 100, threads=1)
    # inside the tolerance of a table everywhere: fine
    T = np.where(np.arange(p.mesh.n_elem) % 2 == 0, 595.0, 1205.0)
    run(dataclasses.replace(p, elem_temperature=T), n=100)


def test_infinite_tolerance_reads_the_nearest_table_however_far():
    # OpenMC at run time: no tolerance check, the nearest table (1200 K) for 5000 K
    p = P.with_temperatures(small_c2(n_particles=600), tolerance=np.inf)
    hot = dataclasses.replace(p, elem_temperature=np.full(p.mesh.n_elem, 5000.0))
    top = dataclasses.replace(p, nuclides=[p.nuclide_temps[q][-1] for q in range(4)], nuclide_temps={},
                              elem_temperature=None)
    assert all(t.temperature == 1200.0 for t in top.nuclides)
    assert_same_histories(run(hot)[1], run(top)[1])
    with pytest.raises(OracleError, match="no cross sections"):
        Oracle(dataclasses.replace(hot, temperature_tolerance=1000.0)).run_batch(1, 10, threads=1)
