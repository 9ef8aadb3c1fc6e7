"""Eigenvalue batches on the GPU against the CPU oracle, through the C ABI.

Per batch: every history's record, the fission bank in (parent, progeny)
order, the resampled source of the next batch and the keff state must be
bit-identical (the per-history keff scores are summed in one fixed order on
both sides); tallies agree to 1e-10 relative.  Mixed meshes, mixed
precision, chunked and sorted schedules give the same histories as the
plain FP64 schedule, and a full fission bank drops sites without faults."""
import numpy as np
import pytest

from macaw_amd import abi, problems as P
from macaw_amd.lib import Context, McwError

from test_gpu_parity import assert_records_equal, assert_tally_close

pytestmark = pytest.mark.gpu

KEFF_FIELDS = ("batch", "fission_sites", "sites_dropped", "n_active", "k_generation", "k_collision",
               "k_absorption", "keff", "keff_std", "keff_used")


def small_k2(n_lattice=10, n_particles=4000):
    return P.k2_tet_eigen(n_lattice=n_lattice, n_particles=n_particles, n_batches=4, n_points=4000)


def assert_sites_equal(g, o):
    assert g.shape == o.shape, (g.shape, o.shape)
    for f in ("parent", "progeny", "element"):
        bad = np.nonzero(g[f] != o[f])[0]
        assert bad.size == 0, f"{f}: {bad.size} sites differ, first {bad[:5]}"
    for f in ("r", "u", "E"):
        a = np.ascontiguousarray(g[f]).view(np.uint64)
        b = np.ascontiguousarray(o[f]).view(np.uint64)
        bad = np.nonzero((a != b).reshape(g.shape[0], -1).any(1))[0]
        assert bad.size == 0, f"{f}: {bad.size} sites differ in bits, first {bad[:5]}"


def run_eigen(problem, batches, n_inactive=1, check_oracle=True, **opt):
    from oracle.pyoracle import Oracle
    n = problem.n_particles
    g = Context(problem, mode=abi.MODE_BATCHED, record_histories=1, run_mode=abi.RUN_EIGENVALUE,
                n_inactive=n_inactive, **opt)
    o = (Oracle(problem, eigenvalue=True, n_inactive=n_inactive, seed_roundtrip=bool(opt.get("seed_roundtrip", 0)))
         if check_oracle else None)
    out = []
    for b in range(1, batches + 1):
        print(f"[eigen] {problem.name} batch {b} n={n} opt={opt}", flush=True)
        sg = g.run_batch(b, n)
        rec = {"stats": sg, "hist": g.histories(n), "keff": g.keff(), "bank": g.fission_bank(),
               "src": g.source_bank()}
        if o is not None:
            so = o.run_batch(b, n, threads=8, record=True)
            rec["o"] = {"stats": so, "hist": o.histories(n), "keff": o.keff(), "bank": o.fission_bank(),
                        "src": o.source_bank()}
        out.append(rec)
    return g, o, out


def test_eigen_parity_tet():
    p = small_k2()
    g, o, res = run_eigen(p, 4, n_inactive=1)
    for b, r in enumerate(res, 1):
        ro = r["o"]
        assert_records_equal(r["hist"], ro["hist"])
        for k in KEFF_FIELDS:
            assert r["keff"][k] == ro["keff"][k], (b, k, r["keff"][k], ro["keff"][k])
        assert r["keff"]["fission_sites"] > 0.3 * p.n_particles
        assert_sites_equal(r["bank"], ro["bank"])
        assert_sites_equal(r["src"], ro["src"])
        assert r["src"].size == p.n_particles
        for k in ("segments", "collisions", "absorptions", "leaks"):
            assert r["stats"][k] == ro["stats"][k], k
    assert_tally_close(g.tally(1), o.tally(1))
    assert_tally_close(g.tally(1, abi.RESULT_SUM_SQ), o.tally(1, abi.RESULT_SUM_SQ))
    # inactive batch 1 scored nothing: SUM is batches 2..4
    assert res[-1]["keff"]["n_active"] == 3


def test_eigen_large_tracklength_beside_small_collision_tally():
    """An eigenvalue batch with the tally pair whose LDS ranges differ: a
    per-element track-length flux over more than 1024 elements (device
    atomics) and a 4-bin universe collision flux (summed in LDS)."""
    p = small_k2(n_particles=3000)
    assert p.mesh.n_elem > 1024
    p.tallies = [
        P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                    scores=[abi.SCORE_FLUX]),
        P.TallySpec(id=2, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_UNIVERSE}],
                    scores=[abi.SCORE_FLUX]),
    ]
    g, o, res = run_eigen(p, 3, n_inactive=1)
    for r in res:
        assert_records_equal(r["hist"], r["o"]["hist"])
        assert r["keff"]["keff"] == r["o"]["keff"]["keff"]
    for tid in (1, 2):
        for kind in (abi.RESULT_SUM, abi.RESULT_SUM_SQ):
            assert_tally_close(g.tally(tid, kind), o.tally(tid, kind))
    assert np.count_nonzero(g.tally(1)) > 0.9 * p.mesh.n_elem


def test_eigen_parity_infinite_medium_hex():
    # reflective HEX8 box: k_absorption is exact per history
    p = P.eigen_box(n_particles=3000)
    g, o, res = run_eigen(p, 3, n_inactive=0)
    for r in res:
        assert_records_equal(r["hist"], r["o"]["hist"])
        for k in KEFF_FIELDS:
            assert r["keff"][k] == r["o"]["keff"][k], k
        assert_sites_equal(r["bank"], r["o"]["bank"])
        assert_sites_equal(r["src"], r["o"]["src"])
        assert abs(r["keff"]["k_absorption"] - 1.0) < 1e-12
    assert_tally_close(g.tally(2), o.tally(2))


def test_eigen_parity_mixed_mesh():
    mesh = P.graded_mixed_box(5, 3, seed=606)
    nucs = [P.flat_nuclide(12.0, 6.0, 0.5), P.flat_nuclide(235.0, 12.0, 4.0, 3.2, nu=(2.5, 1e-8))]
    mats = [P.MaterialData(id=1, nuclides=np.array([0, 1], np.int32), densities=np.array([0.3, 0.05]))]
    p = P.Problem(name="eigen_mixed", mesh=mesh, nuclides=nucs, materials=mats,
                  block_material={int(b): 1 for b in np.unique(mesh.subdomain)}, bcs={0: abi.BC_VACUUM},
                  source=P.volume_watt_source(), tallies=P.flux_tallies(both=True), n_particles=3000,
                  n_batches=3)
    g, o, res = run_eigen(p, 3, n_inactive=1)
    for r in res:
        assert_records_equal(r["hist"], r["o"]["hist"])
        for k in KEFF_FIELDS:
            assert r["keff"][k] == r["o"]["keff"][k], k
        assert_sites_equal(r["bank"], r["o"]["bank"])
        assert_sites_equal(r["src"], r["o"]["src"])
    for tid in (1, 2):
        assert_tally_close(g.tally(tid), o.tally(tid))


def test_eigen_parity_seed_roundtrip_and_trace():
    """MaCaw's Real-stored seeds (rounded through a double each segment) and
    a traced final batch, in an eigenvalue run: still bit-identical."""
    p = small_k2(n_particles=2000)
    g, o, res = run_eigen(p, 3, n_inactive=1, seed_roundtrip=1, trace_histories=50)
    for r in res:
        assert_records_equal(r["hist"], r["o"]["hist"])
        for k in KEFF_FIELDS:
            assert r["keff"][k] == r["o"]["keff"][k], k
        assert_sites_equal(r["bank"], r["o"]["bank"])
        assert_sites_equal(r["src"], r["o"]["src"])
    tr = g.trace()
    assert tr.size > 0 and set(np.unique(tr["history"])) <= set(range(50))


@pytest.mark.parametrize("opt", [{"bank_capacity": 1500}, {"sort_cadence": 1}, {"lanes_min": 1},
                                 {"block_size": 128}])
def test_eigen_schedules_agree(opt):
    p = small_k2(n_particles=4000)
    _, _, ref = run_eigen(p, 3, check_oracle=False)
    _, _, alt = run_eigen(p, 3, check_oracle=False, **opt)
    for a, b in zip(ref, alt):
        assert_records_equal(b["hist"], a["hist"])
        for k in KEFF_FIELDS:
            assert a["keff"][k] == b["keff"][k], k
        assert_sites_equal(b["bank"], a["bank"])
        assert_sites_equal(b["src"], a["src"])


def test_eigen_mixed_precision_close_to_fp64():
    p = small_k2(n_lattice=12, n_particles=20000)
    _, _, f64 = run_eigen(p, 4, check_oracle=False)
    _, _, mix = run_eigen(p, 4, check_oracle=False, precision=abi.PREC_MIXED)
    n = p.n_particles
    a, b = f64[0]["keff"], mix[0]["keff"]
    # batch 1 runs the same histories up to FP32 rounding of segment lengths:
    # a few in 1e3 of them take another path
    assert abs(int(a["fission_sites"]) - int(b["fission_sites"])) <= 1e-2 * n
    for k in ("k_generation", "k_collision", "k_absorption"):
        assert abs(a[k] - b[k]) < 1e-2 * a[k], k
    # later batches start from differently resampled banks: independent
    # estimates of the same keff (spread about 1 / sqrt(n) per batch)
    ka = np.array([r["keff"]["k_generation"] for r in f64[1:]])
    kb = np.array([r["keff"]["k_generation"] for r in mix[1:]])
    assert abs(ka.mean() - kb.mean()) < 6.0 / np.sqrt(n * ka.size)


def test_eigen_full_bank_drops_sites():
    # k_inf = 10: batch 1 (keff 1) asks for about 10 n sites, the bank holds
    # 3 n; batch 2 divides the yield by batch 1's keff (about 10)
    p = P.eigen_box(sig_t=10.0, sig_a=3.0, sig_f=3.0, nu=10.0, n_particles=2000)
    g = Context(p, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE)
    for b in (1, 2):
        g.run_batch(b, p.n_particles)
        k = g.keff()
        if b == 1:
            assert k["fission_sites"] == 3 * p.n_particles and k["sites_dropped"] > 4 * p.n_particles
        else:
            assert k["sites_dropped"] == 0 and abs(k["fission_sites"] - p.n_particles) < 0.2 * p.n_particles
        bank = g.fission_bank()
        key = bank["parent"].astype(np.int64) * (1 << 24) + bank["progeny"]
        assert np.all(np.diff(key) > 0)
        assert g.source_bank().size == p.n_particles


def test_kappa_fission_and_nuclide_bins_match_oracle():
    """kappa-fission scores (track-length and collision, total and per
    nuclide) against the oracle, and kappa-fission = kappa x fission for the
    one fissionable nuclide."""
    from oracle.pyoracle import Oracle
    p = small_k2(n_particles=3000)
    p.tallies = [P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                             scores=[abi.SCORE_FISSION, abi.SCORE_KAPPA_FISSION], nuclides=[4, -1]),
                 P.TallySpec(id=2, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_UNIVERSE}],
                             scores=[abi.SCORE_KAPPA_FISSION])]
    g = Context(p, mode=abi.MODE_BATCHED)
    o = Oracle(p)
    g.run_batch(1, p.n_particles)
    o.run_batch(1, p.n_particles, threads=8)
    for tid in (1, 2):
        assert_tally_close(g.tally(tid), o.tally(tid))
    t = g.tally(1).reshape(-1, 2, 2)  # (cell, nuclide bin, score)
    assert t[:, :, 0].sum() > 0
    np.testing.assert_allclose(t[:, :, 1], 193.4e6 * t[:, :, 0], rtol=1e-12, atol=0)


def test_eigen_errors():
    p = small_k2(n_particles=500)
    g = Context(p, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE)
    with pytest.raises(McwError, match="no eigenvalue batch"):
        g.keff()
    g.run_batch(1, 500)
    with pytest.raises(McwError, match="in order"):
        g.run_batch(3, 500)
    with pytest.raises(McwError, match="in order"):
        g.run_batch(2, 600)
    g.set_options(mode=abi.MODE_PLAIN)
    with pytest.raises(McwError, match="mode BATCHED"):
        g.run_batch(2, 500)
    q = P.eigen_box(sig_f=0.0, n_particles=300)
    h = Context(q, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE)
    with pytest.raises(McwError, match="No fission sites"):
        h.run_batch(1, 300)
