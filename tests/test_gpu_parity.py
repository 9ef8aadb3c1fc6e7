"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Per history the crossed-element sequence (a hash chain over element and event
of every segment), the segment and collision counts, the fate and the final
state (energy, position, tracking seed) must be bit-identical; tallies agree to
1e-10 relative (FP64 atomics change only the summation order)."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context

pytestmark = pytest.mark.gpu

TALLY_RTOL = 1e-10


def small_tet(n_lattice=9, n_materials=4, n_particles=3000):
    mesh = P.tet_cube(n_lattice, seed=101, cache=False)
    mesh.subdomain = P.voronoi_blocks(mesh, n_materials, seed=303)
    return P.Problem(name="tet_small", mesh=mesh, nuclides=P.c2_nuclides(2000, 4),
                     materials=P.c2_materials(n_materials, 4),
                     block_material={b: b + 1 for b in range(n_materials)}, bcs={0: abi.BC_VACUUM},
                     source=P.volume_watt_source(), tallies=P.flux_tallies(both=True),
                     n_particles=n_particles, n_batches=2)


def c1_small(n_particles=1000):
    return P.c1_hex_reflective(n_particles=n_particles, n_batches=2, n_points=20000)


def assert_records_equal(g, o):
    assert g.shape == o.shape
    for f in ("id", "digest", "seed", "n_segments", "n_collisions", "elem", "fate"):
        bad = np.nonzero(g[f] != o[f])[0]
        assert bad.size == 0, f"{f}: {bad.size} histories differ, first {bad[:5]}: {g[f][bad[:3]]} vs {o[f][bad[:3]]}"
    for f in ("E", "x", "y", "z"):
        gb, ob = g[f].view(np.uint64), o[f].view(np.uint64)
        bad = np.nonzero(gb != ob)[0]
        assert bad.size == 0, f"{f}: {bad.size} histories differ in bits, first {bad[:5]}"


def assert_tally_close(a, b, rtol=TALLY_RTOL):
    scale = np.maximum(np.abs(b), 1e-300)
    rel = np.abs(a - b) / scale
    assert rel.max() <= rtol, f"max rel diff {rel.max():.3e} at {rel.argmax()}"


@pytest.fixture(scope="module")
def ctx_probe():
    c = Context()
    yield c
    c.close()


def test_device_math_bit_identical(ctx_probe, oracle_lib):
    po = oracle_lib
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(0, 1, 40000), 2.0 ** -rng.uniform(0, 64, 20000),
                        rng.uniform(0, 1e7, 20000), [1.0, 0.5, 0.25, 1e-300, 2.0**-64]])
    seeds = rng.integers(0, 2**63, x.size, dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(1)
    d = ctx_probe.math_probe(x, seeds)
    lg = np.array([po.log(float(v)) for v in x])
    assert np.array_equal(d["log"].view(np.uint64), lg.view(np.uint64))
    sq = np.sqrt(x)
    assert np.array_equal(d["sqrt"].view(np.uint64), sq.view(np.uint64))
    fr = x - np.floor(x)
    sc = np.array([po.sincos_2pi(float(v)) for v in fr])
    assert np.array_equal(d["sin"].view(np.uint64), sc[:, 0].view(np.uint64))
    assert np.array_equal(d["cos"].view(np.uint64), sc[:, 1].view(np.uint64))
    pr = np.array([po.prn(int(s))[0] for s in seeds])
    assert np.array_equal(d["prn"].view(np.uint64), pr.view(np.uint64))
    fs = np.array([po.future_seed(i * 152917 + 7, int(s)) for i, s in enumerate(seeds)], np.uint64)
    assert np.array_equal(d["future"], fs)


def run_both(problem, mode, batches=(1,), roundtrip=False, **opt):
    from oracle.pyoracle import Oracle
    g = Context(problem, mode=mode, record_histories=1, seed_roundtrip=int(roundtrip), **opt)
    o = Oracle(problem, seed_roundtrip=roundtrip, max_segments=opt.get("max_segments", 0))
    out = []
    for b in batches:
        sg = g.run_batch(b, problem.n_particles)
        so = o.run_batch(b, problem.n_particles, threads=4, record=True)
        out.append((sg, so, g.histories(problem.n_particles), o.histories(problem.n_particles)))
    return g, o, out


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_c1_hex_reflective_parity(mode):
    p = c1_small(1000)
    g, o, res = run_both(p, mode, batches=(1, 2))
    for sg, so, hg, ho in res:
        assert_records_equal(hg, ho)
        for k in ("segments", "collisions", "absorptions", "leaks", "cutoffs", "lost", "reflections"):
            assert sg[k] == so[k], k
        assert sg["lost"] == 0 and sg["leaks"] == 0
        assert abs(sg["track_length"] - so["track_length"]) <= 1e-10 * so["track_length"]
    for tid in (1, 2):
        for kind in (abi.RESULT_VALUE, abi.RESULT_SUM, abi.RESULT_SUM_SQ):
            assert_tally_close(g.tally(tid, kind), o.tally(tid, kind))


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_tet_vacuum_watt_parity(mode):
    p = small_tet()
    g, o, res = run_both(p, mode, batches=(1,))
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    assert sg["segments"] == so["segments"] and sg["leaks"] == so["leaks"]
    assert sg["leaks"] > 0 and sg["lost"] == 0
    for tid in (1, 2):
        assert_tally_close(g.tally(tid), o.tally(tid))


def test_seed_roundtrip_mode_parity():
    """MaCaw's Real-stored seeds (round to double each segment) on both sides."""
    p = c1_small(500)
    g, o, res = run_both(p, abi.MODE_BATCHED, roundtrip=True)
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    assert np.all(hg["seed"] < 2**64)
    # seeds stay representable as doubles
    assert np.all(hg["seed"].astype(np.float64).astype(np.uint64) == hg["seed"])


def test_plain_and_batched_agree_bitwise():
    p = small_tet(n_particles=4000)
    a = Context(p, mode=abi.MODE_PLAIN, record_histories=1)
    b = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    sa = a.run_batch(3, p.n_particles)
    sb = b.run_batch(3, p.n_particles)
    assert_records_equal(a.histories(p.n_particles), b.histories(p.n_particles))
    assert sa["segments"] == sb["segments"]
    assert_tally_close(a.tally(1), b.tally(1))


def test_chunked_bank_matches_single_chunk():
    p = small_tet(n_particles=3000)
    a = Context(p, mode=abi.MODE_PLAIN, record_histories=1, bank_capacity=700)
    b = Context(p, mode=abi.MODE_BATCHED, record_histories=1, bank_capacity=1100)
    c = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    for x in (a, b, c):
        x.run_batch(1, p.n_particles)
    ha, hb, hc = (x.histories(p.n_particles) for x in (a, b, c))
    assert_records_equal(ha, hc)
    assert_records_equal(hb, hc)


def test_tally_reset_and_accumulation():
    p = c1_small(300)
    g = Context(p)
    g.run_batch(1, 300)
    v1 = g.tally(1, abi.RESULT_VALUE).copy()
    g.run_batch(2, 300)
    v2 = g.tally(1, abi.RESULT_VALUE)
    s = g.tally(1, abi.RESULT_SUM)
    ss = g.tally(1, abi.RESULT_SUM_SQ)
    assert_tally_close(s, v1 / 300 + v2 / 300, 1e-13)
    assert_tally_close(ss, (v1 / 300) ** 2 + (v2 / 300) ** 2, 1e-12)
    g.tally_reset()
    assert not g.tally(1).any()


def test_errors_are_reported():
    from macaw_amd.lib import McwError
    p = c1_small(10)
    g = Context(p)
    with pytest.raises(McwError, match="batch numbers start at 1"):
        g.run_batch(0, 10)
    p2 = c1_small(10)
    p2.block_material = {0: 7}
    g2 = Context(p2)
    with pytest.raises(McwError, match="Could not find material 7"):
        g2.run_batch(1, 10)
    p3 = c1_small(10)
    p3.bcs = {b: abi.BC_REFLECT for b in range(5)}
    g3 = Context(p3)
    with pytest.raises(McwError, match="boundary 5"):
        g3.run_batch(1, 10)


def test_union_grid_lookup_matches_per_nuclide_search(oracle_lib):
    """Unionized-grid lookup (LDS bucket table + union index map) gives the
    same bits as the oracle's per-nuclide binary search, including energies
    at grid points, below/above a nuclide's range and between grids."""
    from oracle.pyoracle import Oracle
    nucs = [P.synth_nuclide(3000, 12.0, seed=1),
            P.synth_nuclide(700, 1.0, seed=2, e_min=1e-3, e_max=1e6),
            P.synth_nuclide(50, 56.0, seed=3, e_min=10.0, e_max=2e7)]
    p = c1_small(10)
    p.nuclides = nucs
    p.materials = [P.MaterialData(id=1, nuclides=np.array([0, 1, 2], np.int32), densities=np.array([0.02, 0.03, 0.01]))]
    g = Context(p)
    o = Oracle(p)
    rng = np.random.default_rng(3)
    for q, n in enumerate(nucs):
        E = np.concatenate([np.exp(rng.uniform(np.log(1e-6), np.log(3e7), 20000)), n.energy,
                            0.5 * (n.energy[1:] + n.energy[:-1]), [n.energy[0], n.energy[-1], 1e-7, 1e8],
                            nucs[(q + 1) % 3].energy])
        k, xs = g.xs_probe(q, E)
        want = np.array([o.micro_xs(q, float(e)) for e in E])
        assert np.array_equal(xs.view(np.uint64), want.view(np.uint64)), f"nuclide {q}"


def test_general_tallies_match_oracle():
    """Filters (universe, cell subset, energy, particle), reaction scores and
    nuclide bins, both estimators, laid out as OpenMC's results_ (last filter
    fastest, nuclide-major score bins)."""
    from oracle.pyoracle import Oracle
    p = small_tet(n_particles=2000)
    ne = p.mesh.n_elem
    rng = np.random.default_rng(9)
    cells = np.sort(rng.choice(ne, size=ne // 3, replace=False))[::-1].copy()
    edges = [1e-5, 1.0, 1e3, 1e5, 2e7]
    sc = [abi.SCORE_FLUX, abi.SCORE_TOTAL, abi.SCORE_ABSORPTION, abi.SCORE_SCATTER]
    p.tallies = [
        P.TallySpec(id=11, estimator=abi.EST_TRACKLENGTH,
                    filters=[{"type": abi.FILTER_UNIVERSE}, {"type": abi.FILTER_ENERGY, "edges": edges}],
                    scores=sc, nuclides=[0, 2, -1]),
        P.TallySpec(id=12, estimator=abi.EST_COLLISION,
                    filters=[{"type": abi.FILTER_CELL, "ids": cells}, {"type": abi.FILTER_ENERGY, "edges": edges}],
                    scores=sc, nuclides=[-1]),
        P.TallySpec(id=13, estimator=abi.EST_TRACKLENGTH,
                    filters=[{"type": abi.FILTER_PARTICLE}, {"type": abi.FILTER_CELL}], scores=[abi.SCORE_FLUX]),
        P.TallySpec(id=14, estimator=abi.EST_TRACKLENGTH,
                    filters=[{"type": abi.FILTER_UNIVERSE, "ids": [2, 0]}], scores=[abi.SCORE_TOTAL],
                    nuclides=[1, 3]),
    ]
    for mode in (abi.MODE_PLAIN, abi.MODE_BATCHED):
        g = Context(p, mode=mode)
        o = Oracle(p)
        g.run_batch(1, p.n_particles)
        o.run_batch(1, p.n_particles, threads=2)
        for t in p.tallies:
            a, b = g.tally(t.id), o.tally(t.id)
            assert a.shape == b.shape, t.id
            assert b.any(), t.id
            assert_tally_close(a, b)


def lds_pair_problem(n_particles=3000):
    """A per-element track-length flux over more than 1024 elements (too many
    bins for the persistent kernel's LDS copy) beside a universe-filtered
    collision flux of 4 bins (summed in LDS): the track-length adds must stay
    on the device tally while the collision ones go through LDS."""
    p = small_tet(n_particles=n_particles)
    assert p.mesh.n_elem > 1024
    p.tallies = [
        P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                    scores=[abi.SCORE_FLUX]),
        P.TallySpec(id=2, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_UNIVERSE}],
                    scores=[abi.SCORE_FLUX]),
    ]
    return p


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_large_tracklength_beside_small_collision_tally(mode):
    p = lds_pair_problem()
    g, o, res = run_both(p, mode, batches=(1, 2))
    for sg, so, hg, ho in res:
        assert_records_equal(hg, ho)
    t1, t2 = g.tally(1), g.tally(2)
    assert t1.size == p.mesh.n_elem and t2.size == 4
    assert np.count_nonzero(t1) > 0.9 * t1.size
    for tid in (1, 2):
        for kind in (abi.RESULT_VALUE, abi.RESULT_SUM, abi.RESULT_SUM_SQ):
            assert_tally_close(g.tally(tid, kind), o.tally(tid, kind))
    # the track-length flux totals the track length of the last batch
    sg = res[-1][0]
    assert abs(g.tally(1, abi.RESULT_VALUE).sum() - sg["track_length"]) <= 1e-10 * sg["track_length"]


def test_material_index_by_index_matches_oracle():
    """MaCaw's material index = id - 1 (CollisionKernel.C:110) on a material
    list whose ids are out of order, against the oracle in the same mode."""
    from oracle.pyoracle import Oracle
    from test_oracle_compat import three_material_tet
    p = three_material_tet(n_particles=2000)
    for mode in (abi.MODE_PLAIN, abi.MODE_BATCHED):
        g = Context(p, mode=mode, record_histories=1, material_index=abi.MATERIAL_BY_INDEX)
        o = Oracle(p, material_index=abi.MATERIAL_BY_INDEX)
        g.run_batch(1, p.n_particles)
        o.run_batch(1, p.n_particles, threads=4, record=True)
        assert_records_equal(g.histories(p.n_particles), o.histories(p.n_particles))
        for tid in (1, 2):
            assert_tally_close(g.tally(tid), o.tally(tid))
    p.materials[1].id = 7
    p.block_material = {0: 7, 1: 2, 2: 3}
    from macaw_amd.lib import McwError
    g = Context(p, material_index=abi.MATERIAL_BY_INDEX)
    with pytest.raises(McwError, match="material index 7"):
        g.run_batch(1, 10)


def test_mixed_precision_against_fp64():
    """FP32 exit-side distances with FP64 re-solve of near ties: the exit side
    always equals FP64's choice, so crossed-element sequences agree for all but
    a handful of histories (segment lengths carry FP32 rounding, which can flip
    a collision test d > L that FP64 decides by less than ~1e-7 cm); tallies
    agree far inside their statistical error."""
    mesh = P.tet_cube(14, seed=101, cache=False)
    mesh.subdomain = P.voronoi_blocks(mesh, 4, seed=303)
    p = P.Problem(name="tet_mixed", mesh=mesh, nuclides=P.c2_nuclides(5000, 4), materials=P.c2_materials(4, 4),
                  block_material={b: b + 1 for b in range(4)}, bcs={0: abi.BC_VACUUM},
                  source=P.volume_watt_source(), tallies=P.flux_tallies(both=False), n_particles=20000, n_batches=1)
    a = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    b = Context(p, mode=abi.MODE_BATCHED, record_histories=1, precision=abi.PREC_MIXED)
    sa = a.run_batch(1, p.n_particles)
    sb = b.run_batch(1, p.n_particles)
    ha, hb = a.histories(p.n_particles), b.histories(p.n_particles)
    same = (ha["digest"] == hb["digest"]) & (ha["n_segments"] == hb["n_segments"])
    frac_diff = 1.0 - same.mean()
    assert frac_diff < 1e-3, frac_diff
    dx = np.sqrt((ha["x"] - hb["x"]) ** 2 + (ha["y"] - hb["y"]) ** 2 + (ha["z"] - hb["z"]) ** 2)[same]
    # final positions: a segment's FP32 length is off by ~16u(|c'|+S)/|u.n|
    # (u = 2^-24, |c'| + S ~ 2h, h ~ 0.08 cm here), so ~1e-8 cm for a typical
    # crossing and up to ~1e-5..1e-4 cm for the rare one whose exit side is
    # close to the |u.n| <= 1e-5 re-solve threshold; 20k histories showed a
    # median of 1.6e-9 cm and a maximum of 1.2e-5 cm.
    assert np.median(dx) < 1e-8 and np.percentile(dx, 99) < 1e-6 and dx.max() < 1e-4
    assert np.array_equal(ha["E"][same], hb["E"][same])
    assert sb["fp64_rechecks"] > 0 and sa["fp64_rechecks"] == 0
    assert sb["fp64_rechecks"] < 0.02 * sb["segments"]
    ta, tb = a.tally(1), b.tally(1)
    rel = abs(tb.sum() - ta.sum()) / ta.sum()
    assert rel < 1e-5, rel


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_box_source_parity(mode):
    """Box source partly outside the mesh: sites resampled until inside, the
    element is the lowest-index one containing the point (device grid locator
    against the oracle's brute-force search)."""
    p = small_tet(n_particles=1500)
    p.source = P.box_source((0.6, -0.2, 0.3), (1.3, 0.5, 0.9), E0=1.5e6)
    g, o, res = run_both(p, mode)
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    assert sg["segments"] == so["segments"] and sg["lost"] == 0
    for tid in (1, 2):
        assert_tally_close(g.tally(tid), o.tally(tid))


def test_box_source_missing_the_mesh():
    """A box entirely outside the mesh: every history is lost at birth."""
    from oracle.pyoracle import Oracle
    p = small_tet(n_particles=100)
    p.source = P.box_source((2.0, 2.0, 2.0), (3.0, 3.0, 3.0))
    g, o, res = run_both(p, abi.MODE_BATCHED)
    sg, so, hg, ho = res[0]
    assert sg["lost"] == 100 and so["lost"] == 100 and sg["segments"] == 0
    assert_records_equal(hg, ho)


def test_atomic_rate_probe_adds_every_add(ctx_probe):
    """The atomic-rate probe's adds all land (the array sums to their count)
    and its rate is positive; bad sizes are refused."""
    from macaw_amd.lib import McwError
    for n_bins in (1 << 10, 1 << 20):
        rate, total = ctx_probe.atomic_rate(n_bins)
        assert rate > 0 and total > 0 and total == int(total)
        assert total % (256 * (256 + 8)) == 0  # blocks x 256 threads x (256 + 8 warm-up) adds
    with pytest.raises(McwError, match="power of two"):
        ctx_probe.atomic_rate(1000)


def test_gather_rate_probe(ctx_probe):
    """The gather-rate probe returns a positive rate for arrays inside and
    beyond the L2s; bad sizes are refused."""
    from macaw_amd.lib import McwError
    for n_lines in (1 << 10, 1 << 20):
        assert ctx_probe.gather_rate(n_lines) > 0
    with pytest.raises(McwError, match="power of two"):
        ctx_probe.gather_rate(3000)
