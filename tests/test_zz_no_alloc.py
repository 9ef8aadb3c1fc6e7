"""mcw_run_batch_noalloc through the C ABI against mcw_run_batch.

The no-allocation entry point takes its counters and source bank from the
library's static per-device scratch array; everything else must exist. Two
calls in a row, the second with fewer histories (the array keeps the first
call's contents), must give every history the record mcw_run_batch gives it,
bit for bit, the same event counts, and tallies within 1e-10.  Chunks smaller
than the batch go through the scratch bank one after another; sizes beyond the
array and contexts that would need set-up are refused with a message."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context, McwError

from test_gpu_parity import assert_records_equal, assert_tally_close, small_tet

pytestmark = pytest.mark.gpu

NOALLOC_CHUNK = 986_624  # MCW_NOALLOC_CHUNK


def problem(n):
    p = small_tet(n_particles=n)
    p.tallies = P.flux_tallies(both=True)
    return p


@pytest.mark.parametrize("opt", [dict(), dict(bank_capacity=1700), dict(precision=abi.PREC_MIXED)],
                         ids=["one-chunk", "chunks", "mixed"])
def test_noalloc_equals_run_batch_twice(opt):
    n1, n2 = 6000, 2500
    p = problem(n1)
    ref = Context(p, mode=abi.MODE_BATCHED, record_histories=1, **opt)
    dut = Context(p, mode=abi.MODE_BATCHED, record_histories=1, **opt)
    for c in (ref, dut):  # set-up: allocations happen here, once; both tallies hold batch 1
        c.run_batch(1, n1)
    for batch, n in ((2, n1), (3, n2)):
        print(f"[noalloc] batch {batch}, {n} histories, {opt}", flush=True)
        sr = ref.run_batch(batch, n)
        sd = dut.run_batch_noalloc(batch, n)
        assert_records_equal(dut.histories(n), ref.histories(n))
        for k in ("histories", "segments", "collisions", "absorptions", "leaks", "cutoffs", "lost", "reflections",
                  "fp64_rechecks"):
            assert sd[k] == sr[k], k
        assert abs(sd["track_length"] - sr["track_length"]) <= 1e-10 * sr["track_length"]
        for tid in (1, 2):
            for kind in (abi.RESULT_VALUE, abi.RESULT_SUM, abi.RESULT_SUM_SQ):
                assert_tally_close(dut.tally(tid, kind), ref.tally(tid, kind))
        if "bank_capacity" in opt:
            assert sd["kernel_launches"] == -(-n // opt["bank_capacity"])


def test_noalloc_beyond_one_scratch_chunk():
    """More histories than one scratch bank holds: two chunks through it."""
    n = NOALLOC_CHUNK + 40_000
    p = problem(n)
    ref = Context(p, mode=abi.MODE_BATCHED)
    dut = Context(p, mode=abi.MODE_BATCHED)
    for c in (ref, dut):
        c.run_batch(1, 1000)
    sr = ref.run_batch(2, n)
    sd = dut.run_batch_noalloc(2, n)
    assert sd["kernel_launches"] == 2 and sr["kernel_launches"] == 1
    for k in ("segments", "collisions", "absorptions", "leaks"):
        assert sd[k] == sr[k], k
    assert_tally_close(dut.tally(1), ref.tally(1))


def test_noalloc_refusals():
    p = problem(500)
    c = Context(p, mode=abi.MODE_BATCHED)
    with pytest.raises(McwError, match="already set up"):
        c.run_batch_noalloc(1, 500)  # never set up
    c.run_batch(1, 500)
    c.run_batch_noalloc(2, 500)
    c.set_options(bank_capacity=NOALLOC_CHUNK + 1)
    with pytest.raises(McwError, match="exceeds the scratch bank"):
        c.run_batch_noalloc(3, 500)
    c.set_options(bank_capacity=0, record_histories=1)
    with pytest.raises(McwError, match="records"):
        c.run_batch_noalloc(3, 500)
    c.set_options(record_histories=0, mode=abi.MODE_PLAIN)
    with pytest.raises(McwError, match="BATCHED"):
        c.run_batch_noalloc(3, 500)
    c.set_options(mode=abi.MODE_BATCHED, sort_cadence=1)
    with pytest.raises(McwError, match="sort"):
        c.run_batch_noalloc(3, 500)
    c.set_options(sort_cadence=0)
    c.run_batch_noalloc(3, 500)
