"""The benchmark workloads at full size on the GPU.

C2 is 1,008,162 TET4 with 10^7 histories in one batch; its eigenvalue
variant adds a fissile nuclide.  The oracle cannot run 10^7 histories in test
time, so full size is checked through properties that do not depend on size:

- in batch 1 a history's seed id is its index (the batch offset (b-1) N is
  0), so the first 2e5 histories of the 10^7 batch are the oracle's
  2e5-history batch, record for record, and in an eigenvalue batch 1 their
  fission sites are the oracle's too;
- plain and batched schedules give every one of the 10^7 histories the same
  record;
- every history ends by leaking or being absorbed, and the counts add up;
- each segment adds w l to one element's track-length flux, so the flux sums
  to the batch's total track length; segment and collision counts equal the
  sums of the per-history counts."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context

from test_gpu_eigen import assert_sites_equal
from test_gpu_parity import assert_records_equal

pytestmark = pytest.mark.gpu

N = 10_000_000
N_ORACLE = 200_000


@pytest.fixture(scope="module")
def c2():
    return P.c2_tet_vacuum(n_particles=N)


def test_c2_full_size_prefix_equals_oracle_and_counts_add_up(c2):
    from oracle.pyoracle import Oracle
    print(f"[fullsize] c2 {c2.mesh.n_elem} elements, {N} histories, batched", flush=True)
    g = Context(c2, mode=abi.MODE_BATCHED, record_histories=1)
    st = g.run_batch(1, N)
    h = g.histories(N)
    assert np.array_equal(h["id"], np.arange(1, N + 1))
    # every history ends; counts add up
    assert st["lost"] == 0 and st["leaks"] + st["absorptions"] + st["cutoffs"] == N
    assert np.all(np.isin(h["fate"], [abi.FATE_LEAKED, abi.FATE_ABSORBED, abi.FATE_CUTOFF]))
    assert int(h["n_segments"].sum(dtype=np.int64)) == st["segments"]
    assert int(h["n_collisions"].sum(dtype=np.int64)) == st["collisions"]
    assert np.count_nonzero(h["fate"] == abi.FATE_LEAKED) == st["leaks"]
    # the flux tally (per source particle) holds every segment's w l
    flux = g.tally(1)
    assert flux.size == c2.mesh.n_elem and np.all(flux >= 0)
    assert abs(flux.sum() * N - st["track_length"]) <= 1e-9 * st["track_length"]
    # the first N_ORACLE histories are the oracle's batch of N_ORACLE
    print(f"[fullsize] oracle on the first {N_ORACLE} histories", flush=True)
    o = Oracle(c2)
    so = o.run_batch(1, N_ORACLE, threads=16, record=True)
    assert_records_equal(h[:N_ORACLE], o.histories(N_ORACLE))
    assert int(h["n_segments"][:N_ORACLE].sum(dtype=np.int64)) == so["segments"]


def test_c2_full_size_plain_equals_batched(c2):
    print(f"[fullsize] c2 {N} histories, plain and batched", flush=True)
    a = Context(c2, mode=abi.MODE_PLAIN, record_histories=1)
    b = Context(c2, mode=abi.MODE_BATCHED, record_histories=1)
    sa, sb = a.run_batch(1, N), b.run_batch(1, N)
    assert_records_equal(a.histories(N), b.histories(N))
    for k in ("segments", "collisions", "absorptions", "leaks", "cutoffs"):
        assert sa[k] == sb[k], k
    ta, tb = a.tally(1), b.tally(1)
    rel = np.abs(ta - tb) / np.maximum(np.abs(tb), 1e-300)
    assert rel.max() <= 1e-10


def test_k2_full_size_eigen_batch1_prefix_equals_oracle():
    from oracle.pyoracle import Oracle
    p = P.k2_tet_eigen(n_particles=N)
    print(f"[fullsize] k2 eigenvalue {p.mesh.n_elem} elements, {N} histories", flush=True)
    g = Context(p, mode=abi.MODE_BATCHED, record_histories=1, run_mode=abi.RUN_EIGENVALUE, n_inactive=1)
    st = g.run_batch(1, N)
    k = g.keff()
    assert st["lost"] == 0 and st["leaks"] + st["absorptions"] + st["cutoffs"] == N
    assert k["sites_dropped"] == 0 and 0.5 * N < k["fission_sites"] < 3 * N
    bank = g.fission_bank()
    assert bank.size == k["fission_sites"]
    key = bank["parent"].astype(np.int64) * (1 << 24) + bank["progeny"]
    assert np.all(np.diff(key) > 0)
    src = g.source_bank()
    assert src.size == N and np.all(src["element"] >= 0)
    # batch 1 histories and their fission sites do not depend on N
    n = N_ORACLE // 2
    o = Oracle(p, eigenvalue=True, n_inactive=1)
    o.run_batch(1, n, threads=16, record=True)
    assert_records_equal(g.histories(N)[:n], o.histories(n))
    ob = o.fission_bank()
    assert_sites_equal(bank[bank["parent"] < n], ob)  # parent: the history's index
    # a second batch runs from the resampled bank
    st2 = g.run_batch(2, N)
    k2 = g.keff()
    assert st2["lost"] == 0 and k2["batch"] == 2 and 0.7 < k2["k_generation"] < 1.4
