"""Request server (continuous batching): every request's per-element flux is
the same alone and in a mix (1e-10 relative: only the FP64 atomic summation
order differs), and equals the oracle's track-length flux for that request
(batch 1, the request's particle count, master seed request_id + 1)."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context, Server

pytestmark = pytest.mark.gpu


def meshes():
    t = P.tet_cube(9, seed=101, cache=False)
    t.subdomain = P.voronoi_blocks(t, 4, seed=303)
    tet = P.Problem(name="tet", mesh=t, nuclides=P.c2_nuclides(2000, 4), materials=P.c2_materials(4, 4),
                    block_material={b: b + 1 for b in range(4)}, bcs={0: abi.BC_VACUUM},
                    source=P.volume_watt_source(), tallies=[], n_particles=1, n_batches=1)
    hexp = P.c1_hex_reflective(n_particles=1, n_points=20000)
    hexp.tallies = []
    return [tet, hexp]


def requests():
    out = []
    srcs = [P.point_source((0.4, 0.45, 0.52), 1.5e6), P.box_source((0.1, 0.1, 0.1), (0.6, 0.9, 0.5), 2e6),
            P.volume_watt_source(), P.point_source((0.3, -1.2, 2.0), 2e6), P.box_source((-4, -4, -4), (0, 0, 0))]
    slots = [0, 0, 0, 1, 1]
    sizes = [3000, 12000, 700, 5000, 20000, 1500, 9000, 400]
    rid = 0
    for k, n in enumerate(sizes):
        j = k % len(srcs)
        out.append((rid, slots[j], n, srcs[j]))
        rid += 7
    return out


def contexts(probs):
    return [Context(p) for p in probs]


def test_requests_same_alone_and_in_mix():
    probs = meshes()
    reqs = requests()
    ctx = contexts(probs)
    # alone: one request at a time through its own server
    alone = {}
    for rid, slot, n, src in reqs:
        s = Server(ctx, max_inflight=4)
        s.submit(rid, slot, n, src)
        alone.update(s.run_all())
        s.close()
    # mixed: all at once, small in-flight limit so admission queues them
    s = Server(ctx, max_inflight=3, bank_bytes=1 << 26)
    for rid, slot, n, src in reqs:
        s.submit(rid, slot, n, src)
    mixed = s.run_all()
    s.close()
    assert sorted(alone) == sorted(mixed) == sorted(r[0] for r in reqs)
    for rid, slot, n, src in reqs:
        a, b = alone[rid][1], mixed[rid][1]
        assert a.sum() > 0
        rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-300)
        assert rel.max() <= 1e-10, (rid, rel.max())
        assert alone[rid][0]["segments"] == mixed[rid][0]["segments"]
        assert alone[rid][0]["collisions"] == mixed[rid][0]["collisions"]
        st = mixed[rid][0]
        assert st["t_submit_ms"] <= st["t_admit_ms"] <= st["t_done_ms"]


def test_requests_match_oracle():
    from oracle.pyoracle import Oracle
    probs = meshes()
    reqs = [r for r in requests() if r[2] <= 5000]
    s = Server(contexts(probs), max_inflight=8)
    for rid, slot, n, src in reqs:
        s.submit(rid, slot, n, src)
    got = s.run_all()
    s.close()
    for rid, slot, n, src in reqs:
        p = probs[slot]
        p.source = src
        p.tallies = P.flux_tallies(both=False)
        o = Oracle(p, master_seed=rid + 1)
        so = o.run_batch(1, n, threads=4)
        want = o.tally(1, abi.RESULT_SUM)
        st, flux = got[rid]
        assert st["segments"] == so["segments"], rid
        assert st["collisions"] == so["collisions"], rid
        rel = np.abs(flux - want) / np.maximum(np.abs(want), 1e-300)
        assert rel.max() <= 1e-10, (rid, rel.max())


def test_many_requests_share_rounds():
    """More active requests on one mesh than one round holds: rounds are
    capped at the bank size, and every request completes with the result it
    has alone."""
    probs = meshes()
    ctx = contexts(probs)
    src = P.box_source((0.2, 0.2, 0.2), (0.8, 0.8, 0.8), 1e6)
    reqs = [(1000 + i, 0, 70_000 + 1000 * i, src) for i in range(40)]
    s = Server(ctx, max_inflight=64, bank_bytes=1 << 40)
    for r in reqs:
        s.submit(*r)
    mixed = s.run_all()
    s.close()
    assert len(mixed) == 40
    for rid, slot, n, src_ in (reqs[0], reqs[39]):
        a = Server(ctx, max_inflight=1)
        a.submit(rid, slot, n, src_)
        st, flux = a.run_all()[rid]
        a.close()
        rel = np.abs(flux - mixed[rid][1]) / np.maximum(np.abs(flux), 1e-300)
        assert rel.max() <= 1e-10
        assert st["segments"] == mixed[rid][0]["segments"]
        assert st["histories"] == n


def test_point_source_requests_beyond_the_cache_keys():
    """80 point-source requests of 3,000 histories in one round (the round
    holds up to 2^18): their first segments all add to the source element,
    which the serving kernel's per-wavefront tally cache absorbs, and the
    requests in slices 63 and up, whose keys the cache does not encode, add
    to the device directly.  Every request equals its result alone."""
    probs = meshes()
    ctx = contexts(probs)
    src = P.point_source((0.5, 0.5, 0.5))
    reqs = [(3000 + i, 0, 3_000, src) for i in range(80)]
    s = Server(ctx, max_inflight=96, bank_bytes=1 << 40)
    for r in reqs:
        s.submit(*r)
    mixed = s.run_all()
    s.close()
    assert len(mixed) == 80
    a = Server(ctx, max_inflight=1)
    for rid, slot, n, src_ in reqs:
        a.submit(rid, slot, n, src_)
    alone = a.run_all()
    a.close()
    for rid, slot, n, src_ in reqs:
        st, flux = alone[rid]
        rel = np.abs(flux - mixed[rid][1]) / np.maximum(np.abs(flux), 1e-300)
        assert rel.max() <= 1e-10, (rid, rel.max())
        assert st["segments"] == mixed[rid][0]["segments"]


def test_shortest_remaining_work_first():
    """A large request and small ones admitted together: the small ones run
    in a round of their own (a round holding short requests takes none of a
    long one's bulk), so they finish well before the large one, which still
    completes with its own result."""
    probs = meshes()
    ctx = contexts(probs)
    src = P.box_source((0.2, 0.2, 0.2), (0.8, 0.8, 0.8), 1e6)
    s = Server(ctx, max_inflight=16)
    s.submit(500, 0, 3_000_000, src)
    for i in range(5):
        s.submit(501 + i, 0, 20_000, src)
    got = s.run_all()
    s.close()
    big = got[500][0]
    # the small requests' round holds 1e5 histories, the large request's
    # rounds 3e6: the small ones are done in a fraction of the large one's time
    for i in range(5):
        assert got[501 + i][0]["t_done_ms"] - big["t_admit_ms"] < 0.5 * (big["t_done_ms"] - big["t_admit_ms"]), \
            (got[501 + i][0], big)
    assert big["histories"] == 3_000_000 and big["segments"] > 0
    # the large request alone gives the same result
    a = Server(ctx, max_inflight=1)
    a.submit(500, 0, 3_000_000, src)
    st, flux = a.run_all()[500]
    a.close()
    assert st["segments"] == big["segments"]
    rel = np.abs(flux - got[500][1]) / np.maximum(np.abs(flux), 1e-300)
    assert rel.max() <= 1e-10


def test_server_tune():
    """Round size and aging are set before any request; out-of-range values
    and tuning a busy server are refused; small rounds give the same results."""
    from macaw_amd.lib import McwError
    probs = meshes()
    ctx = contexts(probs)
    src = P.box_source((0.2, 0.2, 0.2), (0.8, 0.8, 0.8), 1e6)
    s = Server(ctx, max_inflight=8)
    with pytest.raises(McwError, match="round_histories"):
        s.tune(100, 10.0)
    with pytest.raises(McwError, match="age_ms"):
        s.tune(1 << 16, -1.0)
    s.tune(5000, 50.0)
    s.submit(1, 0, 23_000, src)
    with pytest.raises(McwError, match="tune before submitting"):
        s.tune(1 << 16, 10.0)
    small = s.run_all()[1]
    s.close()
    a = Server(ctx, max_inflight=8)
    a.submit(1, 0, 23_000, src)
    ref = a.run_all()[1]
    a.close()
    assert small[0]["segments"] == ref[0]["segments"]
    rel = np.abs(small[1] - ref[1]) / np.maximum(np.abs(ref[1]), 1e-300)
    assert rel.max() <= 1e-10
