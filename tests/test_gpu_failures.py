"""Failure handling forced from the host: bad inputs fail at setup with a
message, never on the device; rays the walk cannot continue are counted as
lost and killed (the reference's tolerate_failure) without stopping the
batch; energies below the grid clamp to the grid's first point as the oracle
does; inverted elements change nothing; an over-subscribed serving bank
admits requests one at a time."""
import math

import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context, McwError, Server

from test_gpu_parity import assert_records_equal, assert_tally_close, c1_small, run_both, small_tet

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_source_energy_below_grid_matches_oracle(mode):
    p = c1_small(800)
    p.source = P.point_source((0.3, -0.2, 0.1), 1e-7)   # grid starts at 1e-5 eV
    g, o, res = run_both(p, mode)
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    # reflective box: every history ends by absorption (1/v dominates down there) or cutoff
    assert sg["absorptions"] + sg["cutoffs"] == p.n_particles
    assert sg["absorptions"] == so["absorptions"] and sg["cutoffs"] == so["cutoffs"]


def bad(source_edit):
    p = c1_small(10)
    source_edit(p.source)
    return Context(p)


@pytest.mark.parametrize("edit,msg", [
    (lambda s: setattr(s, "E0", float("nan")), "source energy"),
    (lambda s: setattr(s, "E0", 0.0), "source energy"),
    (lambda s: s.point.__setitem__(0, math.inf), "coordinates must be finite"),
    (lambda s: (setattr(s, "energy", abi.ENERGY_WATT), setattr(s, "watt_a", -1.0)), "Watt"),
    (lambda s: (setattr(s, "angle", abi.ANGLE_LINEAR_MU), s.axis.__setitem__(2, float("nan"))), "unit vector"),
])
def test_bad_sources_fail_at_setup(edit, msg):
    c = bad(edit)
    with pytest.raises(McwError, match=msg):
        c.run_batch(1, 10)


def two_tets(flat):
    """Two tets sharing the side (0, 1, 2); with `flat` the second one's
    fourth node lies in that side's plane (zero volume)."""
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0 if flat else -1]], float)
    conn = np.array([0, 1, 2, 3, 0, 1, 2, 4], np.int32)
    return P.Mesh(xyz=xyz, elem_type=np.full(2, abi.TET4, np.uint8), conn_offsets=np.array([0, 4, 8], np.int64),
                  conn=conn, subdomain=np.zeros(2, np.int32), face_boundary=None)


def test_degenerate_element_fails_at_setup():
    base = c1_small(10)
    for flat in (False, True):
        p = P.Problem(name="two", mesh=two_tets(flat), nuclides=base.nuclides, materials=base.materials,
                      block_material={0: 1}, bcs={0: abi.BC_VACUUM}, source=P.point_source((0.1, 0.1, 0.1)),
                      tallies=[], n_particles=10, n_batches=1)
        if flat:
            with pytest.raises(McwError, match="degenerate element 1"):
                Context(p)
        else:
            Context(p).run_batch(1, 10)


def test_nonfinite_node_fails_at_setup():
    p = c1_small(10)
    p.mesh.xyz = p.mesh.xyz.copy()
    p.mesh.xyz[5, 1] = float("nan")
    with pytest.raises(McwError, match="finite"):
        Context(p)


def test_inverted_elements_change_nothing():
    """Planes are oriented by the centroid and built through each side's three
    lowest node ids, so node order changes no plane; with a point source (a
    volume source samples points from the node order) histories are equal."""
    p = small_tet(n_particles=3000)
    q = small_tet(n_particles=3000)
    p.source = q.source = P.point_source((0.41, 0.43, 0.47), 1.2e6)
    t = q.mesh.conn.reshape(-1, 4).copy()
    t[::2, [1, 2]] = t[::2, [2, 1]]       # every other tet negatively oriented
    q.mesh.conn = np.ascontiguousarray(t.ravel())
    a = Context(p, mode=abi.MODE_BATCHED, record_histories=1)
    b = Context(q, mode=abi.MODE_BATCHED, record_histories=1)
    a.run_batch(1, p.n_particles)
    b.run_batch(1, q.n_particles)
    assert_records_equal(b.histories(q.n_particles), a.histories(p.n_particles))
    assert_tally_close(b.tally(1), a.tally(1))


@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_segment_cap_kills_and_counts(mode):
    p = c1_small(500)
    g, o, res = run_both(p, mode, max_segments=7)
    sg, so, hg, ho = res[0]
    assert_records_equal(hg, ho)
    lost = hg["fate"] == abi.FATE_LOST
    assert sg["lost"] == so["lost"] == lost.sum() > 0
    assert np.all(hg["n_segments"][lost] == 7)


def test_oversubscribed_serving_bank_admits_one_at_a_time():
    p = small_tet(n_particles=1)
    p.tallies = []
    ctx = Context(p)
    reqs = [(i + 1, 0, 4000 + 500 * i, P.box_source((0.1, 0.1, 0.1), (0.9, 0.9, 0.9), 1e6)) for i in range(5)]
    full = Server([ctx], max_inflight=8)
    for r in reqs:
        full.submit(*r)
    want = full.run_all()
    full.close()
    tiny = Server([ctx], max_inflight=8, bank_bytes=1)
    for r in reqs:
        tiny.submit(*r)
    got = tiny.run_all()
    tiny.close()
    for rid, _, n, _ in reqs:
        assert got[rid][0]["segments"] == want[rid][0]["segments"]
        assert_tally_close(got[rid][1], want[rid][1])
    # one at a time: each admission waits for the previous request to finish
    order = sorted(got.values(), key=lambda x: x[0]["t_admit_ms"])
    for prev, nxt in zip(order, order[1:]):
        assert nxt[0]["t_admit_ms"] >= prev[0]["t_done_ms"]
