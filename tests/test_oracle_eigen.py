"""The oracle's eigenvalue mode against what the mathematics fixes.

Infinite medium (reflective cube, one flat fissile nuclide): every history
is absorbed, so the absorption estimator nu sigma_f / sigma_a is exact per
history; with no scattering the collision estimator is exact too; the
track-length and collision estimators and the number of fission sites per
history (k_inf / keff) hold in expectation.  Fission neutrons: Watt energies
(mean 3a/2 + a^2 b/4) and isotropic directions.  The resampling
(OpenMCStudy::synchronizeBanks, src/userobjects/OpenMCStudy.C:454-607) in
its special cases: n banked sites are taken as they are (p = 1), a bank
dividing n is repeated site by site (p = 0); otherwise order is kept, sizes
are exact and the draws are the tracking stream of seed id `batch`.  keff's
average and standard deviation against numpy's mean and standard error of
the active batches' generation values."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P


def _bank(n, seed=3):
    rng = np.random.default_rng(seed)
    b = np.zeros(n, abi.site_dtype())
    b["r"] = rng.uniform(-1, 1, (n, 3))
    b["u"] = rng.normal(size=(n, 3))
    b["E"] = rng.uniform(1, 2e7, n)
    b["parent"] = np.arange(n)
    b["element"] = rng.integers(0, 100, n)
    return b


def test_tree_sum(oracle_lib):
    po = oracle_lib
    rng = np.random.default_rng(1)
    for n in (1, 5, 1023, 1024, 1025, 70000, 3 * 1024 * 1024 + 17):
        v = rng.uniform(0, 1, n)
        import math
        assert abs(po.tree_sum(v) - math.fsum(v)) <= 1e-12 * math.fsum(v)
        iv = rng.integers(0, 1000, n).astype(np.float64)  # exact in any order
        assert po.tree_sum(iv) == iv.sum()


def test_sync_bank_special_cases(oracle_lib):
    po = oracle_lib
    b = _bank(1000)
    # total == n: p = 1, every site once, in order
    out = po.synchronize_bank(b, 2, 1000)
    assert np.array_equal(out, b)
    # n a multiple of total: p = 0, each site n / total times, consecutive
    out = po.synchronize_bank(b[:250], 3, 1000)
    assert np.array_equal(out, np.repeat(b[:250], 4))


@pytest.mark.parametrize("total,n", [(3000, 1000), (1001, 1000), (700, 1000), (333, 1000), (1, 17)])
def test_sync_bank_invariants(oracle_lib, total, n):
    po = oracle_lib
    b = _bank(total, seed=total)
    batch, master = 4, 1
    out = po.synchronize_bank(b, batch, n, master)
    assert out.size == n
    idx = out["parent"]
    base = n // total if total < n else 0
    needed = n % total if total < n else n
    p = needed / total
    # the draws: one per banked site from the tracking stream of seed id `batch`
    seed, draws = po.init_seed(batch, 0, master), []
    for _ in range(total):
        v, seed = po.prn(seed)
        draws.append(v)
    draws = np.array(draws)
    picked = np.repeat(np.arange(total), base + (draws < p).astype(int))
    fin = picked.size
    if fin >= n:
        assert np.array_equal(idx, picked[:n])
    else:
        tail = np.maximum(0, total - (n - fin) + np.arange(n - fin))
        assert np.array_equal(idx, np.concatenate([picked, tail]))
    assert np.all(np.diff(idx[: min(fin, n)]) >= 0)  # bank order kept
    if total < n and fin <= n:  # (a cut removes copies of the last sites)
        assert np.all(np.bincount(idx, minlength=total) >= n // total)


def test_infinite_medium_estimators(oracle_lib):
    prob = P.eigen_box(nu=2.25)
    kinf = 2.25 * 1.2 / 3.0  # nu sigma_f / sigma_a
    o = oracle_lib.Oracle(prob, eigenvalue=True, n_inactive=0)
    N = 20000
    st = o.run_batch(1, N, threads=8)
    k = o.keff()
    assert st["absorptions"] == N
    assert abs(k["k_absorption"] - kinf) <= 1e-12 * kinf  # exact per history
    # track length per history ~ Exp(1/Sigma_a): k_tl per history has sd k_inf
    assert abs(k["k_generation"] - kinf) < 4 * kinf / np.sqrt(N)
    # collisions per history ~ Geometric(sigma_a/sigma_t), score nu sigma_f/sigma_t each
    p = 0.3
    sd_col = 2.25 * 1.2 / 10.0 * np.sqrt(1 - p) / p
    assert abs(k["k_collision"] - kinf) < 4 * sd_col / np.sqrt(N)
    # sites per history: k_inf / keff (keff = 1 in batch 1), Poisson-like spread
    assert abs(k["fission_sites"] / N - kinf) < 4 * np.sqrt(3.0 / N)
    assert k["keff_used"] == 1.0


def test_pure_absorber_is_exact(oracle_lib):
    # sigma_t = sigma_a: one collision per history, collision estimator exact
    prob = P.eigen_box(sig_t=3.0, sig_a=3.0, sig_f=1.5, nu=2.0)
    o = oracle_lib.Oracle(prob, eigenvalue=True)
    o.run_batch(1, 5000, threads=4)
    k = o.keff()
    assert abs(k["k_absorption"] - 1.0) <= 1e-12 and abs(k["k_collision"] - 1.0) <= 1e-12


def test_fission_neutrons_watt_isotropic(oracle_lib):
    prob = P.eigen_box()
    o = oracle_lib.Oracle(prob, eigenvalue=True)
    o.run_batch(1, 40000, threads=8)
    fb = o.fission_bank()
    a, b = 0.988e6, 2.249e-6
    mean = 1.5 * a + 0.25 * a * a * b
    E = fb["E"]
    assert abs(E.mean() - mean) < 4 * E.std() / np.sqrt(E.size)
    u = fb["u"]
    assert np.allclose((u * u).sum(1), 1.0, atol=1e-14)
    assert np.all(np.abs(u.mean(0)) < 4 * np.sqrt(1 / 3 / u.shape[0]))
    assert np.all(np.abs((u * u).mean(0) - 1 / 3) < 4 * np.sqrt(4 / 45 / u.shape[0]))
    # (parent, progeny) order, progeny counting from 0 per parent
    key = fb["parent"].astype(np.int64) * (1 << 24) + fb["progeny"]
    assert np.all(np.diff(key) > 0)
    first = np.r_[True, fb["parent"][1:] != fb["parent"][:-1]]
    assert np.all(fb["progeny"][first] == 0)


def test_generations_keff_and_tallies(oracle_lib):
    prob = P.eigen_box()
    o = oracle_lib.Oracle(prob, eigenvalue=True, n_inactive=2)
    N, kg, used = 3000, [], []
    for b in range(1, 9):
        o.run_batch(b, N, threads=4)
        k = o.keff()
        kg.append(k["k_generation"])
        used.append(k["keff_used"])
        flux = o.tally(1, abi.RESULT_SUM)
        if b <= 2:
            assert k["n_active"] == 0 and k["keff"] == k["k_generation"]
            assert np.all(flux == 0.0)  # inactive batches score nothing
        else:
            assert k["n_active"] == b - 2
            act = np.array(kg[2:])
            assert np.isclose(k["keff"], act.mean(), rtol=1e-14)
            if act.size > 1:
                assert np.isclose(k["keff_std"], act.std(ddof=1) / np.sqrt(act.size), rtol=1e-10)
            assert flux.sum() > 0
        src = o.source_bank()
        assert src.size == N
        # the next batch divides the yield by this batch's keff
    assert used[0] == 1.0 and used[1] == kg[0] and used[2] == kg[1]
    assert abs(np.mean(kg[2:]) - 1.0) < 4 * 1.0 / np.sqrt(N * 6)


def test_eigen_batches_in_order(oracle_lib):
    prob = P.eigen_box()
    o = oracle_lib.Oracle(prob, eigenvalue=True)
    o.run_batch(1, 500, threads=2)
    with pytest.raises(oracle_lib.OracleError, match="in order"):
        o.run_batch(3, 500, threads=2)
    with pytest.raises(oracle_lib.OracleError, match="in order"):
        o.run_batch(2, 400, threads=2)
    o.run_batch(2, 500, threads=2)
    o.run_batch(1, 400, threads=2)  # batch 1 restarts


def test_no_fissionable_material_fails(oracle_lib):
    prob = P.eigen_box(sig_f=0.0)
    o = oracle_lib.Oracle(prob, eigenvalue=True)
    with pytest.raises(oracle_lib.OracleError, match="No fission sites"):
        o.run_batch(1, 200, threads=2)
