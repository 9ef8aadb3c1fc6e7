"""Staged per-element flux tally (option tally_stage) on TET4 meshes.

In a BATCHED fixed-source launch whose only track-length tally is a plain
per-element flux of more than 1024 bins, the persistent kernel stages each
segment's score in LDS by bucket of 2^14 bins and writes full buckets to a
record pool; k_stage_reduce sums each bucket in LDS and adds the sums to the
tally.  Only the order of the FP64 additions changes, so:

- histories are the same record for record as with one device atomic per
  segment (the staging kernels also read a coarser energy-bucket table, which
  must not change a lookup);
- every element's flux agrees with the atomic path within 1e-12 and with the
  oracle within the 1e-10 of the other parity tests;
- a pool too small for the launch sends the rest to device atomics with the
  same result, and so do launches split by bank_capacity;
- a pool sized from a small batch grows for a larger one, and SUM and SUM_SQ
  over several batches agree with the atomic path."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context

from test_gpu_parity import assert_records_equal, assert_tally_close

pytestmark = pytest.mark.gpu

N = 200_000
N_ORACLE = 20_000


@pytest.fixture(scope="module")
def c2s():
    p = P.c2_tet_vacuum(n_lattice=30, n_particles=N, tallies=P.flux_tallies(both=True))
    assert p.mesh.n_elem > 4 * (1 << 14)  # several buckets
    return p


def run(p, n, **opt):
    g = Context(p, mode=abi.MODE_BATCHED, **opt)
    st = g.run_batch(1, n)
    out = (st, g.histories(n) if opt.get("record_histories") else None, g.tally(1, abi.RESULT_VALUE).copy(),
           g.tally(2, abi.RESULT_VALUE).copy())
    g.close()
    return out


@pytest.fixture(scope="module")
def atomic_fp64(c2s):
    print(f"[stage] {c2s.mesh.n_elem} elements, {N} histories, device atomics", flush=True)
    return run(c2s, N, tally_stage=-1, record_histories=1)


def same_counts(a, b):
    for k in ("segments", "collisions", "absorptions", "leaks", "lost"):
        assert a[k] == b[k], k


def test_staged_equals_atomics_with_records(c2s, atomic_fp64):
    st0, h0, t0, c0 = atomic_fp64
    assert st0["stage_ms"] == 0.0
    print("[stage] staged, records", flush=True)
    st, h, t, c = run(c2s, N, record_histories=1)
    assert st["stage_ms"] > 0.0
    same_counts(st, st0)
    assert_records_equal(h, h0)
    assert_tally_close(t, t0, 1e-12)
    assert np.array_equal(c, c0) or np.abs(c - c0).max() <= 1e-12 * np.abs(c0).max()
    assert abs(t.sum() - st["track_length"]) <= 1e-9 * st["track_length"]


def test_staged_without_records(c2s, atomic_fp64):
    st0, _, t0, _ = atomic_fp64
    print("[stage] staged, no records", flush=True)
    st, _, t, _ = run(c2s, N)
    assert st["stage_ms"] > 0.0
    same_counts(st, st0)
    assert_tally_close(t, t0, 1e-12)


def test_staged_pool_overflow_and_split_launches(c2s, atomic_fp64):
    st0, _, t0, _ = atomic_fp64
    print("[stage] staged, 3000-record pool", flush=True)
    st, _, t, _ = run(c2s, N, tally_stage=3000)
    same_counts(st, st0)
    assert_tally_close(t, t0, 1e-12)
    print("[stage] staged, launches of 30000 histories", flush=True)
    st, _, t, _ = run(c2s, N, bank_capacity=30_000, sort_cadence=1)
    assert st["kernel_launches"] == 7 and st["stage_ms"] > 0.0
    same_counts(st, st0)
    assert_tally_close(t, t0, 1e-12)


def test_staged_mixed_precision_equals_atomic_mixed(c2s):
    print("[stage] mixed precision, staged and atomics", flush=True)
    a = run(c2s, N, precision=abi.PREC_MIXED, tally_stage=-1, record_histories=1)
    b = run(c2s, N, precision=abi.PREC_MIXED)
    same_counts(b[0], a[0])
    assert b[0]["fp64_rechecks"] == a[0]["fp64_rechecks"]
    assert_tally_close(b[2], a[2], 1e-12)


def test_staged_matches_oracle(c2s):
    from oracle.pyoracle import Oracle
    print(f"[stage] staged against the oracle, {N_ORACLE} histories", flush=True)
    st, h, t, c = run(c2s, N_ORACLE, record_histories=1)
    assert st["stage_ms"] > 0.0
    o = Oracle(c2s)
    so = o.run_batch(1, N_ORACLE, threads=16, record=True)
    assert st["segments"] == so["segments"]
    assert_records_equal(h, o.histories(N_ORACLE))
    assert_tally_close(t, o.tally(1, abi.RESULT_VALUE))
    assert_tally_close(c, o.tally(2, abi.RESULT_VALUE))


def test_staged_pool_grows_and_sums_accumulate(c2s):
    """Batches of 20,000 then 200,000 then 50,000 histories in one context:
    the pool sized from the first batch grows for the second (the slices of a
    launch are sized from the last batch's segments per history), and the
    SUM over three batches equals the device-atomic context's within 1e-12."""
    print("[stage] growing batches, staged and atomics", flush=True)
    sizes = (20_000, 200_000, 50_000)
    ctx = {ts: Context(c2s, mode=abi.MODE_BATCHED, tally_stage=ts) for ts in (0, -1)}
    for b, n in enumerate(sizes, start=1):
        st = {ts: g.run_batch(b, n) for ts, g in ctx.items()}
        assert st[0]["stage_ms"] > 0.0 and st[-1]["stage_ms"] == 0.0
        same_counts(st[0], st[-1])
    for kind in (abi.RESULT_SUM, abi.RESULT_SUM_SQ):
        assert_tally_close(ctx[0].tally(1, kind), ctx[-1].tally(1, kind), 1e-12)
    for g in ctx.values():
        g.close()


def test_staged_cell_filtered_flux(c2s):
    """A track-length flux over a cell filter of every third element, listed
    in reverse: its bins are the filter's order, not element ids, and their
    number (55,030 on this mesh) is not a multiple of the 2^14-bin buckets.
    Staged, it equals the device-atomic launch within 1e-12 and the oracle
    within 1e-10."""
    from oracle.pyoracle import Oracle
    cells = np.arange(c2s.mesh.n_elem, dtype=np.int64)[::3][::-1].copy()
    p = P.Problem(name="cells", mesh=c2s.mesh, nuclides=c2s.nuclides, materials=c2s.materials,
                  block_material=c2s.block_material, bcs=c2s.bcs, source=c2s.source,
                  tallies=[P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH,
                                       filters=[{"type": abi.FILTER_CELL, "ids": cells}],
                                       scores=[abi.SCORE_FLUX])],
                  n_particles=N_ORACLE, n_batches=1)
    print(f"[stage] cell filter of {cells.size} elements", flush=True)
    out = {}
    for ts in (0, -1):
        g = Context(p, mode=abi.MODE_BATCHED, tally_stage=ts)
        st = g.run_batch(1, N_ORACLE)
        out[ts] = (st, g.tally(1, abi.RESULT_VALUE).copy())
        g.close()
    assert out[0][0]["stage_ms"] > 0.0 and out[-1][0]["stage_ms"] == 0.0
    assert out[0][1].size == cells.size
    assert_tally_close(out[0][1], out[-1][1], 1e-12)
    o = Oracle(p)
    o.run_batch(1, N_ORACLE, threads=16)
    assert_tally_close(out[0][1], o.tally(1, abi.RESULT_VALUE))
