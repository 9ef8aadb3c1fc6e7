"""Mixed TET4/HEX8 meshes on the GPU: a graded box whose refined corner is
Kuhn tets meeting the hexes through split interfaces, with a source linear in
mu.  FP64 histories are bit-identical to the oracle's in every schedule;
mixed precision (FP32 records for the tets, FP64 for the hexes) changes the
element sequence of fewer than 1e-3 of the histories."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context

from test_gpu_parity import assert_records_equal, assert_tally_close, run_both

pytestmark = pytest.mark.gpu


def small_mixed(n_particles=4000, source=None):
    mesh = P.graded_mixed_box(8, 5, seed=11)
    mesh.subdomain = P.voronoi_blocks(mesh, 8, seed=12)
    src = source or P.linear_mu(P.volume_watt_source(), (0.3, -0.5, 0.8), 0.5)
    return P.Problem(name="mixed_small", mesh=mesh, nuclides=P.c2_nuclides(3000, 8, jitter=0.4),
                     materials=P.c2_materials(8, 8, seed=13), block_material={b: b + 1 for b in range(8)},
                     bcs={0: abi.BC_VACUUM}, source=src, tallies=P.flux_tallies(both=True),
                     n_particles=n_particles, n_batches=1)


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_mixed_mesh_volume_source_parity(mode):
    p = small_mixed()
    g, o, res = run_both(p, mode, batches=(1, 2))
    for sg, so, hg, ho in res:
        assert_records_equal(hg, ho)
        assert sg["segments"] == so["segments"] and sg["collisions"] == so["collisions"]
        assert sg["hex_segments"] == so["hex_segments"] and 0 < sg["hex_segments"] < sg["segments"]
        assert sg["lost"] == 0 and sg["leaks"] > 0
    for tid in (1, 2):
        assert_tally_close(g.tally(tid), o.tally(tid))


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_mixed_mesh_point_source_in_tet_corner(mode):
    src = P.linear_mu(P.point_source((0.05, 0.04, 0.06), 1.5e6), (1.0, 1.0, 1.0), -0.7)
    p = small_mixed(n_particles=3000, source=src)
    g, o, res = run_both(p, mode, batches=(1,))
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    assert sg["segments"] == so["segments"]
    assert_tally_close(g.tally(1), o.tally(1))


@pytest.mark.parametrize("opt", [dict(mode=abi.MODE_BATCHED, sort_cadence=1),
                                 dict(mode=abi.MODE_PLAIN, use_graph=1, sort_cadence=8)],
                         ids=["batched-sorted", "plain-graph-sorted"])
def test_mixed_mesh_schedules(opt):
    p = small_mixed(n_particles=3000)
    opt = dict(opt)
    mode = opt.pop("mode")
    g, o, res = run_both(p, mode, batches=(1,), **opt)
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    assert_tally_close(g.tally(1), o.tally(1))


def test_mixed_mesh_mixed_precision():
    """Histories start among the tets of the refined corner: their FP32 exit
    distances leave final positions different from FP64's in the low bits
    (the sign that the FP32 path ran) while the element sequences agree."""
    src = P.linear_mu(P.point_source((0.05, 0.04, 0.06), 1.5e6), (1.0, 1.0, 1.0), 0.7)
    p = small_mixed(n_particles=20000, source=src)
    a = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    b = Context(p, mode=abi.MODE_BATCHED, record_histories=1, precision=abi.PREC_MIXED)
    sa, sb = a.run_batch(1, p.n_particles), b.run_batch(1, p.n_particles)
    ha, hb = a.histories(p.n_particles), b.histories(p.n_particles)
    same = (ha["digest"] == hb["digest"]) & (ha["n_segments"] == hb["n_segments"])
    assert 1.0 - same.mean() < 1e-3
    dx = np.sqrt((ha["x"] - hb["x"]) ** 2 + (ha["y"] - hb["y"]) ** 2 + (ha["z"] - hb["z"]) ** 2)[same]
    assert np.mean(dx > 0) > 0.5 and np.median(dx) < 1e-8 and dx.max() < 1e-4
    assert sa["fp64_rechecks"] == 0
    ta, tb = a.tally(1), b.tally(1)
    assert abs(tb.sum() - ta.sum()) / ta.sum() < 1e-5
    c = Context(p, mode=abi.MODE_PLAIN, record_histories=1, precision=abi.PREC_MIXED)
    c.run_batch(1, p.n_particles)
    # the plain path has no FP32 variant: it stays FP64
    assert_records_equal(c.histories(p.n_particles), ha)


def test_mixed_mesh_serving_matches_oracle():
    from macaw_amd.lib import Server
    from oracle.pyoracle import Oracle
    p = small_mixed(n_particles=1)
    srcs = [P.linear_mu(P.box_source((0.02, 0.02, 0.02), (0.5, 0.4, 0.3), 2e6), (0, 0, 1), 0.8),
            P.point_source((0.7, 0.2, 0.9), 1e6)]
    ctx = Context(p)
    s = Server([ctx], max_inflight=4)
    for rid, src in enumerate(srcs):
        s.submit(rid + 40, 0, 2500, src)
    got = s.run_all()
    s.close()
    for rid, src in enumerate(srcs):
        p.source = src
        p.tallies = P.flux_tallies(both=False)
        o = Oracle(p, master_seed=rid + 41)
        so = o.run_batch(1, 2500, threads=4)
        st, flux = got[rid + 40]
        assert st["segments"] == so["segments"]
        assert_tally_close(flux, o.tally(1, abi.RESULT_SUM))


def test_hex_mesh_parity_and_precision():
    """An all-HEX8 mesh runs its own FP64 kernel (6 planes at a 6-slot stride,
    the neighbour record apart) in both precisions: histories equal the
    oracle's, and mixed precision changes nothing.  (The mixed-mesh layout
    with FP32 hex records measured slower in FP64 on the extruded assembly,
    185.6 against 172.0 ms per batch, and only 2% faster in mixed precision.)"""
    from oracle.pyoracle import Oracle
    from test_gpu_parity import assert_records_equal, assert_tally_close
    p = P.c1_hex_reflective(n_particles=4000, n_batches=1, n_points=20000)
    g = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    o = Oracle(p)
    sg = g.run_batch(1, p.n_particles)
    o.run_batch(1, p.n_particles, threads=8, record=True)
    assert_records_equal(g.histories(p.n_particles), o.histories(p.n_particles))
    assert_tally_close(g.tally(1), o.tally(1))
    assert sg["hex_segments"] == sg["segments"]
    m = Context(p, mode=abi.MODE_BATCHED, record_histories=1, precision=abi.PREC_MIXED)
    sm = m.run_batch(1, p.n_particles)
    assert_records_equal(m.histories(p.n_particles), g.histories(p.n_particles))
    assert sm["fp64_rechecks"] == 0 and sm["segments"] == sg["segments"]
