"""Element temperatures on the GPU against the oracle, through the C ABI.

The library resolves OpenMC's "nearest" table per element and nuclide once,
on the host, into device materials and nuclide instances; the oracle looks
the table up per segment.  Histories must agree to the bit, tallies to
1e-10 relative, including tallies with nuclide bins (which see the caller's
nuclide, whichever table it was read at)."""
import dataclasses

import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context, McwError

from test_gpu_eigen import KEFF_FIELDS, assert_sites_equal
from test_gpu_parity import assert_records_equal, assert_tally_close

pytestmark = pytest.mark.gpu


def nuclide_tallies():
    return [P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                        scores=[abi.SCORE_FLUX, abi.SCORE_TOTAL, abi.SCORE_ABSORPTION]),
            P.TallySpec(id=2, estimator=abi.EST_COLLISION,
                        filters=[{"type": abi.FILTER_ENERGY, "edges": [1e-5, 1.0, 1e3, 1e5, 2e7]}],
                        scores=[abi.SCORE_TOTAL, abi.SCORE_SCATTER], nuclides=[0, 3, -1])]


def tet_temps(n_lattice=10, n_particles=4000, axis=0):
    mesh = P.tet_cube(n_lattice, seed=101, cache=False)
    mesh.subdomain = P.voronoi_blocks(mesh, 4, seed=303)
    p = P.Problem(name="tet_temps", mesh=mesh, nuclides=P.c2_nuclides(3000, 4), materials=P.c2_materials(4, 4),
                  block_material={b: b + 1 for b in range(4)}, bcs={0: abi.BC_VACUUM},
                  source=P.volume_watt_source(), tallies=nuclide_tallies(), n_particles=n_particles,
                  n_batches=2)
    return P.with_temperatures(p, axis=axis)


def mixed_temps(n_particles=3000):
    mesh = P.graded_mixed_box(5, 3, seed=606)
    mesh.subdomain = P.voronoi_blocks(mesh, 4, seed=303)
    p = P.Problem(name="mixed_temps", mesh=mesh, nuclides=P.c2_nuclides(3000, 4), materials=P.c2_materials(4, 4),
                  block_material={b: b + 1 for b in range(4)}, bcs={0: abi.BC_VACUUM},
                  source=P.volume_watt_source(), tallies=nuclide_tallies(), n_particles=n_particles,
                  n_batches=1)
    return P.with_temperatures(p, axis=1)


def run_both(p, batches=(1,), **opt):
    from oracle.pyoracle import Oracle
    g = Context(p, record_histories=1, **opt)
    o = Oracle(p)
    n = p.n_particles
    for b in batches:
        print(f"[temperature] {p.name} batch {b} n={n} opt={opt}", flush=True)
        sg = g.run_batch(b, n)
        so = o.run_batch(b, n, threads=8, record=True)
        assert_records_equal(g.histories(n), o.histories(n))
        for k in ("segments", "collisions", "absorptions", "leaks"):
            assert sg[k] == so[k], k
    for t in p.tallies:
        assert_tally_close(g.tally(t.id), o.tally(t.id))
    return g, o


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_temperature_parity_tet(mode):
    g, o = run_both(tet_temps(), batches=(1, 2), mode=mode)
    assert g.tally(2).reshape(4, 3, 2)[:, :2].sum() > 0  # nuclide bins scored


def test_temperature_parity_mixed_mesh():
    run_both(mixed_temps(), mode=abi.MODE_BATCHED)


def test_temperature_sorted_schedule():
    # (material, element) sorting over the device materials temperatures make
    run_both(tet_temps(n_particles=3000, axis=2), mode=abi.MODE_BATCHED, sort_cadence=1)


def test_temperature_mixed_precision_close_to_fp64():
    p = tet_temps(n_lattice=14, n_particles=20000)
    a = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    b = Context(p, mode=abi.MODE_BATCHED, record_histories=1, precision=abi.PREC_MIXED)
    a.run_batch(1, p.n_particles)
    b.run_batch(1, p.n_particles)
    ha, hb = a.histories(p.n_particles), b.histories(p.n_particles)
    same = (ha["digest"] == hb["digest"]) & (ha["n_segments"] == hb["n_segments"])
    assert 1.0 - same.mean() < 1e-3
    ta, tb = a.tally(1), b.tally(1)
    assert abs(tb.sum() - ta.sum()) / ta.sum() < 1e-5


def test_temperature_eigen_parity():
    from oracle.pyoracle import Oracle
    p = P.with_temperatures(P.k2_tet_eigen(n_lattice=10, n_particles=3000, n_points=3000), axis=2)
    g = Context(p, mode=abi.MODE_BATCHED, record_histories=1, run_mode=abi.RUN_EIGENVALUE, n_inactive=1)
    o = Oracle(p, eigenvalue=True, n_inactive=1)
    n = p.n_particles
    for b in (1, 2, 3):
        print(f"[temperature] eigen batch {b}", flush=True)
        g.run_batch(b, n)
        o.run_batch(b, n, threads=8, record=True)
        assert_records_equal(g.histories(n), o.histories(n))
        kg, ko = g.keff(), o.keff()
        for k in KEFF_FIELDS:
            assert kg[k] == ko[k], (b, k)
        assert_sites_equal(g.fission_bank(), o.fission_bank())
        assert_sites_equal(g.source_bank(), o.source_bank())
    assert_tally_close(g.tally(1), o.tally(1))


def test_temperature_set_and_clear():
    from oracle.pyoracle import Oracle
    p = tet_temps(n_particles=2000)
    g = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    # cleared: every element reads the first tables, as a problem with no temperatures
    g.set_temperatures(None)
    g.run_batch(1, p.n_particles)
    o = Oracle(dataclasses.replace(p, elem_temperature=None))
    o.run_batch(1, p.n_particles, threads=8, record=True)
    assert_records_equal(g.histories(p.n_particles), o.histories(p.n_particles))
    # uniform 1190 K: the 1200 K tables everywhere
    g.tally_reset()
    g.set_temperatures(np.full(p.mesh.n_elem, 1190.0), tolerance=10.0)
    g.run_batch(1, p.n_particles)
    o = Oracle(dataclasses.replace(p, elem_temperature=np.full(p.mesh.n_elem, 1190.0), temperature_tolerance=10.0))
    o.run_batch(1, p.n_particles, threads=8, record=True)
    assert_records_equal(g.histories(p.n_particles), o.histories(p.n_particles))


def test_temperature_errors():
    p = tet_temps(n_particles=500)
    g = Context(p, mode=abi.MODE_BATCHED)
    g.set_temperatures(p.elem_temperature, tolerance=10.0)
    with pytest.raises(McwError, match="no cross sections"):
        g.run_batch(1, 500)
    g.set_temperatures(p.elem_temperature[:-1], tolerance=200.0)
    with pytest.raises(McwError, match="one temperature per element"):
        g.run_batch(1, 500)
    with pytest.raises(McwError, match="finite"):
        g.set_temperatures(np.full(p.mesh.n_elem, np.nan))
    with pytest.raises(McwError, match="at this temperature"):
        g.add_temperature(0, p.nuclide_temps[0][0])
    with pytest.raises(McwError, match="no such nuclide"):
        g.add_temperature(9, p.nuclide_temps[0][0])
    fis = P.fissile(p.nuclide_temps[0][0], 0.5)
    with pytest.raises(McwError, match="fission data"):
        g.add_temperature(0, dataclasses.replace(fis, temperature=1500.0))
    g.set_temperatures(p.elem_temperature, tolerance=200.0)
    g.run_batch(1, 500)


def test_temperature_random_field_many_device_materials():
    """Random element temperatures over the four tables of every nuclide:
    each material appears at each of the four table choices, sixteen device
    materials scattered element by element; plain, batched and (material,
    element)-sorted schedules all equal the oracle."""
    from oracle.pyoracle import Oracle
    p = tet_temps(n_lattice=10, n_particles=3000)
    rng = np.random.default_rng(77)
    p = dataclasses.replace(p, elem_temperature=rng.uniform(250.0, 1300.0, p.mesh.n_elem))
    o = Oracle(p)
    o.run_batch(1, p.n_particles, threads=8, record=True)
    ho = o.histories(p.n_particles)
    for opt in ({"mode": abi.MODE_PLAIN}, {"mode": abi.MODE_BATCHED}, {"mode": abi.MODE_BATCHED, "sort_cadence": 1}):
        print(f"[temperature] random field {opt}", flush=True)
        g = Context(p, record_histories=1, **opt)
        g.run_batch(1, p.n_particles)
        assert_records_equal(g.histories(p.n_particles), ho)
        for t in p.tallies:
            assert_tally_close(g.tally(t.id), o.tally(t.id))
