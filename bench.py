#!/usr/bin/env python3
"""Benchmark: histories/s and segments/s of the particle-tracking hot path.

Workloads (one "step" = one batch of fixed-source histories):
  c2 (default) unit-cube TET4 mesh from a Delaunay triangulation of a
     jittered 54^3 lattice (1,008,162 tets), vacuum boundaries, 4 Voronoi
     materials of 4 synthetic nuclides (50k-point grids), uniform volumetric
     Watt source, 1e7 histories per batch, per-element track-length flux;
  c3 the same generator at 8M tets (107^3 lattice), 16 materials of 4 of 32
     nuclides on the unionized grid, 1e8 histories per batch in one launch;
  c5 a 20M-element mixed TET4/HEX8 mesh graded toward a corner (tets meet
     hexes through split interfaces), 128 materials of 4 of 64 nuclides with
     100k-point grids, low densities, a volumetric Watt source linear in mu,
     1e7 histories per batch.
Inputs are synthetic and resident in HBM before the timed region.

`value` is the batched (persistent, fused) FP64 path, whose histories are
bit-identical to the CPU oracle's.  Reported beside it: the mixed-precision
form (FP32 exit distances with FP64 re-solve of near ties) with its tally
difference from the FP64 run, the plain event-based GPU path on the same
batch, the dominant kernel's roofline (algorithmic bytes per launch over its
HIP-event duration), the serving mix (c2 runs only: 2000 requests over
resident 100k / 1M / 4M-tet meshes, Poisson arrivals at 70% of measured
capacity, p50/p99 latency), an eigenvalue run (c2 runs only: C2's mesh with
a fissile nuclide, 1e7 histories per generation, 2 inactive + 3 active
batches, each with its fission-bank sort and resampling), the reference's
BEAVRS assembly example run from its driver input (c2 runs only: the
example's Exodus mesh extruded to 95,504 HEX8, eigenvalue, 1e7 histories per
batch), the 8M-tet batching + precision workload (c2 runs only: 16
materials of 4 of 32 nuclides, 1e8 histories per batch in one launch,
sorted sources; FP64, mixed precision and plain), the 20M-element stress
workload (c2 runs only: mixed TET4/HEX8, 64 nuclides, 1e7 histories per
batch; FP64, mixed precision and plain) and a CPU baseline
(the FP64 oracle, std::thread on the host cores, on a bounded sample of the
same histories).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c3|c5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "histories/sec + segments/sec at 1M tets vs plain GPU; p50/p99 serving latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats, n_nuc_per_mat, b_seg=276.0):
    """92 B per history (bank record), B_seg per segment (TET4 FP64: particle
    read 76 + write 36, element record 148, tally 16), 64*n_nuc + 64 per
    collision (two bracketing points of 4 tables per nuclide + state)."""
    return 92.0 * stats["histories"] + b_seg * stats["segments"] + (64.0 * n_nuc_per_mat + 64.0) * stats["collisions"]


def read_traffic(workload):
    """Bytes per launch of the dominant kernel between L2 and memory, from the
    committed rocprofv3 --pmc summary (per-size TCC_EA0 read and write
    request counts and memory-side atomics, times their request sizes)."""
    path = os.path.join(ROOT, "profiles", f"k_persistent_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split("=")[1])
            except ValueError:
                return None
    return None


def cpu_baseline(problem, n_hist_full, seconds=12.0):
    from oracle.pyoracle import Oracle
    cores = os.cpu_count() or 1
    o = Oracle(problem)
    pilot = 40000
    t = time.perf_counter()
    o.run_batch(1, pilot, threads=cores)
    dt = time.perf_counter() - t
    n = int(min(max(pilot * seconds / max(dt, 1e-3), 10000), 2_000_000))
    log(f"cpu baseline: pilot {pilot} histories in {dt:.2f}s -> sample {n} histories on {cores} threads")
    o2 = Oracle(problem)
    t = time.perf_counter()
    st = o2.run_batch(1, n, threads=cores)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "histories/s", "cores": cores, "kind": "port",
            "sample": f"batch 1, histories 1..{n} of the same workload (seed ids (b-1)*N+i with N={n}), "
                      f"{st['segments']} segments",
            "segments_per_s": st["segments"] / dt, "seconds": dt}


def bench_serving(args, tuning):
    """Serving mix: resident 100k / 1M / 4M-tet meshes, requests with point or
    box sources and particle counts log-uniform in [1e4, 1e7], Poisson
    arrivals at `--serve-load` of the capacity measured by a saturated run."""
    from macaw_amd import abi, problems as P
    from macaw_amd.lib import Context, Server
    from macaw_amd.serving import run_closed_loop, run_open_loop, summarize

    t = time.perf_counter()
    probs = P.c4_meshes()
    log(f"serving meshes {[p.mesh.n_elem for p in probs]} generated ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    ctxs = [Context(p, mode=abi.MODE_BATCHED, lanes_min=tuning.get("lanes_min", 0)) for p in probs]
    bank = int(args.serve_bank_gib * (1 << 30))
    srv = Server(ctxs, max_inflight=args.serve_inflight, bank_bytes=bank)
    srv.tune(args.serve_round, args.serve_age_ms)
    log(f"server up ({time.perf_counter() - t:.1f}s)")
    warm = P.c4_requests(n=6, seed=504, n_max=1e5)
    for r in warm:
        r.request_id += 2_000_000
    run_closed_loop(srv, warm)
    cal = P.c4_requests(n=args.serve_calibrate, seed=506)
    for r in cal:
        r.request_id += 1_000_000
    t = time.perf_counter()
    run_closed_loop(srv, cal)
    cap = sum(r.n_particles for r in cal) / (time.perf_counter() - t)
    reqs = P.c4_requests(n=args.serve_requests, seed=505)
    mean_n = sum(r.n_particles for r in reqs) / len(reqs)
    rate = args.serve_load * cap / mean_n
    reqs = P.c4_requests(n=args.serve_requests, seed=505, rate=rate)
    log(f"serving: capacity {cap:.3e} histories/s (saturated, {len(cal)} requests); "
        f"{len(reqs)} requests at {rate:.2f}/s ({args.serve_load:.0%} load, mean {mean_n:.3e} particles)")
    t = time.perf_counter()
    res = run_open_loop(srv, reqs, log=log)
    wall = time.perf_counter() - t
    srv.close()
    for c in ctxs:
        c.close()
    out = summarize(res, wall)
    out.update({"rate_requests_per_s": rate, "offered_load": args.serve_load, "capacity_histories_per_s": cap,
                "meshes_tets": [p.mesh.n_elem for p in probs], "max_inflight": args.serve_inflight,
                "bank_gib": args.serve_bank_gib, "round_histories": args.serve_round,
                "age_ms": args.serve_age_ms, "particles": "log-uniform [1e4, 1e7]",
                "sources": "point or box (1:1), Watt", "tally": "per-element track-length flux per request",
                "arrivals": "Poisson, seed 505", "precision": "fp64"})
    log(f"serving: p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, "
        f"{out['histories_per_s']:.3e} histories/s over {wall:.1f}s")
    return out


def tune_eigen_lanes(prob, n, candidates=(16, 24, 32, 40), log_name="eigen", tally_stage=0):
    """lanes_min for an eigenvalue kernel, picked on the workload itself: each
    candidate runs batches 1 and 2 (the second from a fission source) on a
    fresh context; the fastest batch-2 transport wins.  Results do not depend
    on lanes_min, only the time does."""
    from macaw_amd import abi
    from macaw_amd.lib import Context
    best, best_ms = candidates[0], float("inf")
    for lm in candidates:
        ctx = Context(prob, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE, n_inactive=0, lanes_min=lm,
                      tally_stage=tally_stage)
        ctx.run_batch(1, n)
        ms = ctx.run_batch(2, n)["time_ms"]
        ctx.close()
        log(f"{log_name}: lanes_min {lm}: batch 2 transport {ms:.1f} ms")
        if ms < best_ms:
            best, best_ms = lm, ms
    return best


def bench_eigen(args, tuning):
    """Eigenvalue leg: C2's mesh and materials with a fissile nuclide added
    (keff near 1), n histories per generation, 2 inactive + 3 active
    batches; each batch is transport plus the batch close (keff reduction,
    fission-bank sort, resampling into the next source)."""
    from macaw_amd import abi, problems as P
    from macaw_amd.lib import Context

    n = args.eigen_histories
    t = time.perf_counter()
    prob = P.k2_tet_eigen(n_particles=n)
    stage = tuning.get("tally_stage", args.tally_stage or 0)
    lanes = args.lanes_min if args.no_autotune and args.lanes_min else tune_eigen_lanes(prob, n, tally_stage=stage)
    ctx = Context(prob, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE, n_inactive=2,
                  lanes_min=lanes, block_size=tuning.get("block_size", 0), tally_stage=stage)
    log(f"eigen: {prob.name} ready ({time.perf_counter() - t:.1f}s)")
    ctx.run_batch(1, n)  # warm: allocations and code objects
    rows = []
    t = time.perf_counter()
    for b in range(1, 6):
        st = ctx.run_batch(b, n)
        k = ctx.keff()
        rows.append((st, k))
        log(f"eigen batch {b}: {st['time_ms']:.1f} ms transport + {k['close_ms']:.1f} ms close, "
            f"{st['segments'] / n:.1f} segments and {st['collisions'] / n:.2f} collisions per history, "
            f"{k['fission_sites']} sites, k_gen {k['k_generation']:.5f}, keff {k['keff']:.5f} +- {k['keff_std']:.5f}")
    wall = time.perf_counter() - t
    ctx.close()
    segs = sum(st["segments"] for st, _ in rows)
    k = rows[-1][1]
    return {"histories_per_s": 5 * n / wall, "segments_per_s": segs / wall, "ms_per_batch": wall / 5 * 1e3,
            "transport_ms_per_batch": sum(st["time_ms"] for st, _ in rows) / 5,
            "close_ms_per_batch": sum(kk["close_ms"] for _, kk in rows) / 5,
            "keff": k["keff"], "keff_std": k["keff_std"], "k_generation": [kk["k_generation"] for _, kk in rows],
            "fission_sites_per_history": sum(kk["fission_sites"] for _, kk in rows) / (5 * n),
            "segments_per_history": segs / (5 * n), "histories_per_batch": n, "batches": 5, "inactive": 2,
            # the source moves from the volumetric start toward the fundamental
            # mode, so the work per history changes from batch to batch
            "per_batch": [{"transport_ms": st["time_ms"], "kernel_ms": st["kernel_ms"],
                           "segments_per_history": st["segments"] / n,
                           "collisions_per_history": st["collisions"] / n,
                           "ns_per_segment": st["kernel_ms"] * 1e6 / max(st["segments"], 1)} for st, _ in rows],
            "lanes_min": lanes,
            "workload": f"k2_tet4_{prob.mesh.n_elem // 1000}k_eigen",
            "description": "C2 mesh and materials plus a fissile nuclide (A=235, 85% of absorption is fission, "
                           "nu 2.43, Watt spectrum), vacuum, batch 1 from a volumetric Watt source"}


def bench_assembly(args):
    """The reference's BEAVRS 3.1% assembly example as its driver input sets
    it up (tests/inputs/beavrs_assembly_3d.i: the example's Exodus mesh
    extruded to 95,504 HEX8, reflective, fission source on z = 100, synthetic
    stand-ins for its materials), eigenvalue with n histories per batch: one
    warm batch, then 4 batches timed."""
    import subprocess
    import tempfile
    from macaw_amd import abi, build, problems as P
    from macaw_amd.lib import Context

    n = args.assembly_histories
    inp = os.path.join(ROOT, "tests", "inputs", "beavrs_assembly_3d.i")
    with tempfile.TemporaryDirectory() as d:
        dump = os.path.join(d, "assembly.dump")
        subprocess.run([build.DRIVER, "-i", inp, "--dump-problem", dump], check=True, stdout=subprocess.DEVNULL)
        prob = P.load_dump(dump)
    prob.tallies = [P.TallySpec(id=1, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_CELL}],
                                scores=[abi.SCORE_KAPPA_FISSION])]
    lanes = args.lanes_min if args.no_autotune and args.lanes_min else tune_eigen_lanes(prob, n, log_name="assembly")
    ctx = Context(prob, mode=abi.MODE_BATCHED, run_mode=abi.RUN_EIGENVALUE, n_inactive=0, lanes_min=lanes)
    ctx.run_batch(1, n)
    rows = []
    t = time.perf_counter()
    for b in range(2, 6):
        rows.append(ctx.run_batch(b, n))
    wall = time.perf_counter() - t
    k = ctx.keff()
    ctx.close()
    segs = sum(st["segments"] for st in rows)
    log(f"assembly: {wall / 4 * 1e3:.1f} ms per batch of {n}, keff {k['keff']:.5f} +- {k['keff_std']:.5f}")
    return {"histories_per_s": 4 * n / wall, "segments_per_s": segs / wall, "ms_per_batch": wall / 4 * 1e3,
            "segments_per_history": segs / (4 * n),
            "collisions_per_history": sum(st["collisions"] for st in rows) / (4 * n),
            "keff": k["keff"], "keff_std": k["keff_std"], "histories_per_batch": n, "batches_timed": 4,
            "lanes_min": lanes,
            "workload": f"beavrs_assembly_hex8_{prob.mesh.n_elem // 1000}k_eigen",
            "description": "the reference's examples/beavrs_assembly mesh (17x17 pin cells, QUAD4) extruded to 4 "
                           "layers of HEX8 over 400 cm, reflective, fission source on z = 100, synthetic nuclides"}


def bench_large(args, tuning, which):
    """A large workload at full size beside the C2 line, sources sorted by
    (material, element), one warm batch then timed batches in FP64 and in
    mixed precision, the mixed form's tally error against FP64 over the same
    batches, and one plain (event-based) batch of up to 1e7 histories for the
    ratio.
      c3: the batching + precision workload, 8,066,911 TET4 (107^3 lattice),
          16 materials of 4 of 32 nuclides on a 1.6M-point union grid, 1e8
          histories per batch in one launch;
      c5: the stress workload, 19,999,360 mixed TET4/HEX8 elements graded
          toward a corner, 128 materials of 4 of 64 nuclides on a 6.4M-point
          union grid, low densities, sources linear in mu, 1e7 histories per
          batch."""
    from macaw_amd import abi, problems as P
    from macaw_amd.lib import Context

    # lanes_min per precision, from sweeps on these workloads (one launch per
    # batch): C3 FP64 24 (1331 ms; 16-40 within 1.5%), mixed 32 (1154 against
    # 1168 ms at 24); C5 48 for both (FP64 157.0 against 157.9 ms at 40, mixed
    # 126.6 against 128.8 ms)
    if which == "c3":
        n, n_steps, lanes = args.c3_histories, args.c3_steps, (args.c3_lanes_min, args.c3_lanes_min_mixed)
        gen = P.c3_tet_vacuum
    else:
        n, n_steps, lanes = args.c5_histories, args.c5_steps, (args.c5_lanes_min, args.c5_lanes_min)
        gen = P.c5_mixed
    t = time.perf_counter()
    prob = gen(n_particles=n)
    all_tet = bool((prob.mesh.elem_type == abi.TET4).all())
    log(f"{which}: {prob.mesh.n_elem} elements, {len(prob.materials)} materials, {len(prob.nuclides)} nuclides "
        f"({time.perf_counter() - t:.1f}s)")
    # a whole batch in one launch (up to 2^27 histories): C3's 1e8 histories in
    # one launch took 1365 ms per batch against 1399 ms in six of 2^24 (mixed
    # precision 1202 against 1242 ms), since each launch ends on a tail of
    # long histories
    bank = min(n, 1 << 27)
    opts = dict(mode=abi.MODE_BATCHED, sort_cadence=1, block_size=tuning.get("block_size", 0), bank_capacity=bank)
    kind = "tet4" if all_tet else "mixed"
    out = {"workload": f"{which}_{kind}_{prob.mesh.n_elem // 1000}k_vacuum", "n_elem": prob.mesh.n_elem,
           "histories_per_step": n, "materials": len(prob.materials), "nuclides": len(prob.nuclides),
           "schedule": f"persistent, sorted sources, lanes_min {lanes[0]} (FP64) and {lanes[1]} (mixed), "
                       f"{-(-n // bank)} launch(es) of up to {bank} histories"}
    ctxs = {}
    ns = 4 if all_tet else 0
    # the two other ceilings of these kernels, measured on arrays of the
    # workload's size: random 128-byte line reads (the element records, about
    # 160 B per element) and FP64 atomic adds (the flux tally)
    glines = 1 << min(25, max(16, (prob.mesh.n_elem * 160 // 128 - 1).bit_length()))
    abins = 1 << min(28, max(10, (prob.mesh.n_elem - 1).bit_length()))
    for form, prec, lm in (("fp64", abi.PREC_FP64, lanes[0]), ("mixed", abi.PREC_MIXED, lanes[1])):
        ctx = Context(prob, precision=prec, lanes_min=lm, **opts)
        ctxs[form] = ctx
        steps, wall = timed_steps(ctx, n, n_steps, 1, f"{which} {form}")
        segs = sum(s["segments"] for s in steps)
        kern_s = sum(s["kernel_ms"] for s in steps) * 1e-3
        if form == "fp64":
            roof = roofline_of(steps, n, 4, 276.0)
            grate, (arate, _) = ctx.gather_rate(glines), ctx.atomic_rate(abins)
        else:
            roof = roofline_of(steps, n, 4, 212.0, b_hex=252.0)
        roof["kernel"] = f"k_persistent<{ns}>" if form == "fp64" else f"k_persistent<{ns}, mixed>"
        # element lines a segment reads: FP64 the plane record and the
        # neighbour record (a hex's 256-byte block), mixed precision one record
        lps = 2.0 if form == "fp64" else 1.0
        out[form] = {"histories_per_s": n * n_steps / wall, "segments_per_s": segs / wall,
                     "ms_per_step": wall / n_steps * 1e3, "segments_per_history": segs / (n * n_steps),
                     "hex_segment_fraction": sum(s["hex_segments"] for s in steps) / max(1, segs),
                     "fp64_rechecks_per_segment": sum(s["fp64_rechecks"] for s in steps) / max(1, segs),
                     "kernel_launches_per_step": steps[0]["kernel_launches"], "roofline": roof,
                     "gather_roofline": {"bound": "random 128-byte line reads", "achieved": lps * segs / kern_s,
                                         "peak": grate, "unit": "lines/s", "frac": lps * segs / kern_s / grate,
                                         "lines_per_segment": lps, "array_lines": glines,
                                         "note": "peak measured by mcw_debug_gather_rate"},
                     "atomic_roofline": {"bound": "memory-side FP64 atomics", "achieved": segs / kern_s,
                                         "peak": arate, "unit": "atomic adds/s", "frac": segs / kern_s / arate,
                                         "array_bins": abins,
                                         "note": "one tally add per segment; peak measured by mcw_debug_atomic_rate"}}
        log(f"{which} {form}: {out[form]['histories_per_s']:.3e} histories/s, frac {roof['frac']:.3f}, "
            f"gather {out[form]['gather_roofline']['frac']:.3f}, atomics {out[form]['atomic_roofline']['frac']:.3f}")
    out["mixed"]["tally_error_vs_fp64"] = tally_error(ctxs["fp64"], ctxs["mixed"], n_steps + 1)
    for c in ctxs.values():
        c.close()
    pctx = Context(prob, mode=abi.MODE_PLAIN)
    pn = min(n, 10_000_000)
    pctx.run_batch(200, pn)
    t = time.perf_counter()
    pst = pctx.run_batch(1, pn)
    pw = time.perf_counter() - t
    pctx.close()
    out["plain_gpu"] = {"histories_per_s": pn / pw, "segments_per_s": pst["segments"] / pw, "histories": pn,
                        "ms_per_step": pw * 1e3}
    out["speedup_vs_plain"] = out["fp64"]["histories_per_s"] / out["plain_gpu"]["histories_per_s"]
    out["mixed_speedup_vs_plain"] = out["mixed"]["histories_per_s"] / out["plain_gpu"]["histories_per_s"]
    log(f"{which} plain: {out['plain_gpu']['histories_per_s']:.3e} histories/s; batched "
        f"{out['speedup_vs_plain']:.2f}x, mixed {out['mixed_speedup_vs_plain']:.2f}x")
    return out


WORKLOADS = {
    # name: (generator, default histories per step, description)
    "c2": ("c2_tet_vacuum", 10_000_000, "1M TET4 unit cube, vacuum, 4 materials / 4 nuclides x 50k points"),
    "c3": ("c3_tet_vacuum", 100_000_000, "8M TET4 unit cube, vacuum, 16 materials / 32 nuclides x 50k points "
                                          "(jittered grids, 1.6M-point union), chunked bank"),
    "c5": ("c5_mixed", 10_000_000, "20M mixed TET4/HEX8 unit cube graded toward a corner, vacuum, 128 materials / "
                                   "64 nuclides x 100k points (6.4M-point union), low densities, volumetric Watt "
                                   "source linear in mu"),
}


def tally_error(fp64_ctx, mixed_ctx, n_batches):
    """Per-element flux of the mixed-precision run against the FP64 run over
    the same batches: relative differences and differences in units of the
    FP64 estimate's standard error (from SUM and SUM_SQ over the batches)."""
    import numpy as np
    from macaw_amd import abi
    f, m = fp64_ctx.tally(1, abi.RESULT_SUM), mixed_ctx.tally(1, abi.RESULT_SUM)
    f2 = fp64_ctx.tally(1, abi.RESULT_SUM_SQ)
    nz = f > 0
    rel = np.abs(m[nz] - f[nz]) / f[nz]
    mean = f / n_batches
    var = np.maximum(f2 / n_batches - mean * mean, 0.0) / max(n_batches - 1, 1)
    sig = np.sqrt(var)
    ok = sig > 0
    z = np.abs(m[ok] - f[ok]) / n_batches / sig[ok]
    # elements whose standard error is itself well estimated (relative
    # standard error below 10%): with few batches an element can show a
    # near-zero sample variance by chance
    good = ok & (sig < 0.1 * np.abs(mean))
    zg = np.abs(m[good] - f[good]) / n_batches / sig[good]
    # The sample variance of one element over n batches has n - 1 degrees of
    # freedom ((n-1) s^2 / sigma^2 ~ chi^2_{n-1}): over millions of elements
    # a few show s a thousand times below sigma by chance, and their z reads
    # in the tens whatever the difference.  The pooled figure takes each
    # element's variance as the median sample variance of the elements with
    # the nearest mean flux (groups of ~2^15), over the median of
    # chi^2_{n-1} / (n-1).
    zp = None
    if n_batches >= 2 and ok.any():
        from scipy.stats import chi2
        order = np.argsort(mean)
        pooled = np.empty_like(var)
        for g in np.array_split(order, max(1, f.size >> 15)):
            pooled[g] = np.median(var[g])
        pooled /= chi2.ppf(0.5, n_batches - 1) / (n_batches - 1)
        okp = pooled > 0
        zp = np.abs(m[okp] - f[okp]) / n_batches / np.sqrt(pooled[okp])
    # one batch gives no standard error: the sigma figures need two or more
    return {"total_rel": float(abs(m.sum() - f.sum()) / f.sum()), "elem_rel_max": float(rel.max()),
            "elem_rel_median": float(np.median(rel)), "elem_rel_p99": float(np.percentile(rel, 99)),
            "max_sigma": float(z.max()) if z.size else None,
            "frac_beyond_3sigma": float((z > 3).mean()) if z.size else None,
            "max_sigma_rse_below_0.1": float(zg.max()) if zg.size else None,
            "elements_rse_below_0.1": int(good.sum()),
            "max_sigma_pooled": float(zp.max()) if zp is not None and zp.size else None,
            "frac_beyond_3sigma_pooled": float((zp > 3).mean()) if zp is not None and zp.size else None,
            "batches": n_batches}


def timed_steps(ctx, N, steps, warmup, label):
    for w in range(warmup):
        st = ctx.run_batch(100 + w, N)
        log(f"{label} warmup {w}: {st['time_ms']:.1f} ms, {st['segments']} segments")
    out = []
    t = time.perf_counter()
    for k in range(steps):
        out.append(ctx.run_batch(k + 1, N))
    wall = time.perf_counter() - t
    for k, st in enumerate(out):
        log(f"{label} step {k}: {st['time_ms']:.1f} ms device, {st['segments']} segments, "
            f"{st['collisions']} collisions, leaks {st['leaks']}, lost {st['lost']}, rechecks {st['fp64_rechecks']}")
    return out, wall


def roofline_of(steps, N, n_nuc, b_seg, b_hex=348.0):
    """B_seg applies to TET4 segments, b_hex to HEX8 segments: 348 B with FP64
    planes (particle 76 + 36, six planes 192 + six neighbours and material 28,
    tally 16), 252 B from the FP32 record (planes 96 + neighbours and
    material 28)."""
    segs = sum(s["segments"] for s in steps)
    hexs = sum(s["hex_segments"] for s in steps)
    colls = sum(s["collisions"] for s in steps)
    kern_ms = sum(s["kernel_ms"] for s in steps)
    stage_ms = sum(s["stage_ms"] for s in steps)
    launches = sum(s["kernel_launches"] for s in steps)
    b_alg = algorithmic_bytes({"histories": N * len(steps), "segments": segs - hexs, "collisions": colls},
                              n_nuc, b_seg) + b_hex * hexs
    per_launch_bytes = b_alg / launches
    per_launch_s = kern_ms / launches / 1e3
    achieved = per_launch_bytes / per_launch_s / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_launch": per_launch_bytes,
           "launch_ms": per_launch_s * 1e3, "launches_per_step": launches / len(steps)}
    if stage_ms > 0:
        # staged flux tally: launch_ms is the transport kernel plus the
        # kernels that sum the staged records (k_stage_count, k_stage_scatter,
        # k_stage_reduce); the algorithmic bytes are the same
        out["transport_ms"] = (kern_ms - stage_ms) / launches
        out["stage_ms"] = stage_ms / launches
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--histories", type=int, default=0, help="histories per step (default: the workload's)")
    ap.add_argument("--lattice", type=int, default=0, help="Delaunay lattice points per side (default: the workload's)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lanes-min", type=int, default=0)
    ap.add_argument("--sort-cadence", type=int, default=0, help="sort sources by (material, element) without autotune")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tally-stage", type=int, default=None,
                    help="tally_stage option of the C2 contexts (0: staged flux tally, -1: device atomics; "
                         "default: the autotuner's choice, 0 without it)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving mix")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-calibrate", type=int, default=200, help="requests in the saturated capacity run")
    ap.add_argument("--serve-load", type=float, default=0.7, help="offered load as a fraction of capacity")
    ap.add_argument("--serve-inflight", type=int, default=32)
    ap.add_argument("--serve-bank-gib", type=float, default=4.0)
    ap.add_argument("--serve-round", type=int, default=1 << 18, help="histories per serving round")
    ap.add_argument("--serve-age-ms", type=float, default=250.0,
                    help="age past which a request is served before shorter ones")
    ap.add_argument("--no-eigen", action="store_true", help="skip the eigenvalue leg")
    ap.add_argument("--eigen-histories", type=int, default=10_000_000)
    ap.add_argument("--no-assembly", action="store_true", help="skip the BEAVRS assembly leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 8M-tet batching + precision leg")
    ap.add_argument("--c3-histories", type=int, default=100_000_000)
    ap.add_argument("--c3-steps", type=int, default=2)
    ap.add_argument("--c3-lanes-min", type=int, default=24)
    ap.add_argument("--c3-lanes-min-mixed", type=int, default=32)
    ap.add_argument("--no-c5", action="store_true", help="skip the 20M-element stress leg")
    ap.add_argument("--c5-histories", type=int, default=10_000_000)
    ap.add_argument("--c5-steps", type=int, default=3)
    ap.add_argument("--c5-lanes-min", type=int, default=48)
    ap.add_argument("--assembly-histories", type=int, default=10_000_000)
    ap.add_argument("--gather-sweep", action="store_true",
                    help="only measure the device's rate of pseudo-random 128-byte line reads (k_gather_probe) "
                         "on arrays of 2^16 .. 2^25 lines")
    ap.add_argument("--atomic-sweep", action="store_true",
                    help="only measure the device's FP64 atomic-add rate (k_atomic_probe) on arrays of "
                         "2^10 .. 2^25 doubles")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(histories shard by id range across GPUs with one tally all-reduce per batch; "
              f"that path is not built)", file=sys.stderr)
        sys.exit(2)

    from macaw_amd import abi, problems as P
    from macaw_amd.lib import Context

    if args.gather_sweep:
        c = Context()
        rows = []
        for lg in (16, 19, 20, 21, 22, 23, 24, 25):
            rate = c.gather_rate(1 << lg)
            rows.append({"lines": 1 << lg, "mbytes": (128 << lg) / 1e6, "lines_per_s": rate,
                         "GB_per_s": rate * 128 / 1e9})
            log(f"gather probe: 2^{lg} lines ({(128 << lg) / 1e6:.0f} MB): {rate:.3e} lines/s, {rate * 128 / 1e9:.0f} GB/s")
        print(json.dumps({"gather_sweep": "k_gather_probe", "rows": rows}))
        return
    if args.atomic_sweep:
        c = Context()
        rows = []
        for lg in (10, 13, 16, 19, 20, 22, 23, 25):
            rate, total = c.atomic_rate(1 << lg)
            rows.append({"bins": 1 << lg, "mbytes": (8 << lg) / 1e6, "adds_per_s": rate, "array_sum": total})
            log(f"atomic probe: 2^{lg} bins ({(8 << lg) / 1e6:.2f} MB): {rate:.3e} adds/s, array sum {total:.0f}")
        print(json.dumps({"atomic_sweep": "k_atomic_probe", "rows": rows}))
        return

    gen_name, n_default, desc = WORKLOADS[args.workload]
    N = args.histories or n_default
    t0 = time.perf_counter()
    gen_kw = {"n_particles": N}
    if args.lattice:
        gen_kw["n_lattice"] = args.lattice
    prob = getattr(P, gen_name)(**gen_kw)
    n_elem = prob.mesh.n_elem
    n_mat, n_nuc = len(prob.materials), len(prob.nuclides)
    nuc_per_mat = max(len(m.nuclides) for m in prob.materials)
    all_tet = bool((prob.mesh.elem_type == abi.TET4).all())
    element = "TET4" if all_tet else "TET4+HEX8"
    workload = f"{args.workload}_{'tet4' if all_tet else 'mixed'}_{n_elem // 1000}k_vacuum"
    log(f"{workload}: {n_elem} tets, {prob.mesh.n_nodes} nodes, {n_mat} materials, {n_nuc} nuclides "
        f"({time.perf_counter() - t0:.1f}s)")

    ctx = Context(prob, mode=abi.MODE_BATCHED, block_size=args.block, lanes_min=args.lanes_min,
                  sort_cadence=args.sort_cadence, tally_stage=args.tally_stage or 0)
    tuning = {"block_size": ctx.opts.block_size, "lanes_min": ctx.opts.lanes_min,
              "waves_per_eu": ctx.opts.waves_per_eu, "sort_cadence": ctx.opts.sort_cadence,
              "tally_stage": ctx.opts.tally_stage}
    if not args.no_autotune:
        from macaw_amd.autotune import autotune
        t = time.perf_counter()
        tuning, rows = autotune(ctx, min(N, 10_000_000), log=log)
        if args.tally_stage is not None:
            tuning["tally_stage"] = args.tally_stage
        ctx.set_options(**tuning)
        log(f"autotune: best {tuning} ({time.perf_counter() - t:.1f}s)")
    ctx.tally_reset()  # tallies then hold exactly the warmup + timed batches
    steps, wall = timed_steps(ctx, N, args.steps, args.warmup, "fp64")
    segs = sum(s["segments"] for s in steps)
    colls = sum(s["collisions"] for s in steps)
    hist_s = N * args.steps / wall
    seg_s = segs / wall
    roof = roofline_of(steps, N, nuc_per_mat, 276.0)
    kname = "k_persistent<4>" if all_tet else "k_persistent<0>"
    if "stage_ms" in roof:
        kname += " + staged-tally sums"
    roof.update({"traffic": read_traffic(workload), "kernel": kname})
    # the per-element flux tally makes one FP64 atomic add per segment, and
    # the device performs every FP64 atomic at the memory side: the kernel's
    # other ceiling is the device's atomic rate, measured here on an array
    # of the tally's size
    # (with the staged tally the transport kernel makes no device atomic per
    # segment: its scores go to LDS and a record pool, summed after it)
    staged = sum(s["stage_ms"] for s in steps) > 0
    transport_s = sum(s["kernel_ms"] - s["stage_ms"] for s in steps) * 1e-3
    seg_per_launch_s = segs / transport_s
    atomic_roof = None
    if not staged:
        bins = 1 << max(10, (n_elem - 1).bit_length())
        arate, _ = ctx.atomic_rate(bins)
        atomic_roof = {"bound": "memory-side FP64 atomics", "achieved": seg_per_launch_s, "peak": arate,
                       "unit": "atomic adds/s", "frac": seg_per_launch_s / arate, "array_bins": bins,
                       "note": "one tally add per segment (kernel time); peak measured by mcw_debug_atomic_rate"}
        log(f"atomics: {seg_per_launch_s:.3e} adds/s in the kernel against {arate:.3e} measured ceiling "
            f"({seg_per_launch_s / arate:.2f})")
    # and random element-record lines: an FP64 TET4 segment reads the
    # 128-byte plane record and the line holding its neighbour record
    glines = 1 << min(25, max(16, (n_elem * 160 // 128 - 1).bit_length()))
    grate = ctx.gather_rate(glines)
    gather_roof = {"bound": "random 128-byte line reads", "achieved": 2.0 * seg_per_launch_s, "peak": grate,
                   "unit": "lines/s", "frac": 2.0 * seg_per_launch_s / grate, "lines_per_segment": 2.0,
                   "array_lines": glines,
                   "note": "transport kernel time; peak measured by mcw_debug_gather_rate"}

    mixed = None
    if not args.no_mixed:
        mctx = Context(prob, mode=abi.MODE_BATCHED, precision=abi.PREC_MIXED)
        mctx.set_options(**tuning)
        msteps, mwall = timed_steps(mctx, N, args.steps, args.warmup, "mixed")
        mroof = roofline_of(msteps, N, nuc_per_mat, 212.0, b_hex=252.0)
        mroof["kernel"] = ("k_persistent<4, mixed>" if all_tet else "k_persistent<0, mixed>") + \
            (" + staged-tally sums" if "stage_ms" in mroof else "")
        mixed = {"histories_per_s": N * args.steps / mwall,
                 "segments_per_s": sum(s["segments"] for s in msteps) / mwall,
                 "ms_per_step": mwall / args.steps * 1e3,
                 "fp64_rechecks_per_segment": sum(s["fp64_rechecks"] for s in msteps)
                 / max(1, sum(s["segments"] for s in msteps)),
                 "roofline": mroof,
                 "tally_error_vs_fp64": tally_error(ctx, mctx, args.steps + args.warmup)}
        log(f"mixed: {mixed['histories_per_s']:.3e} histories/s; tally vs fp64 {mixed['tally_error_vs_fp64']}")
        mctx.close()

    plain = None
    if not args.no_plain:
        pctx = Context(prob, mode=abi.MODE_PLAIN)
        pN = min(N, 10_000_000)
        pctx.run_batch(200, pN)  # warm
        t = time.perf_counter()
        pst = pctx.run_batch(1, pN)
        pw = time.perf_counter() - t
        plain = {"histories_per_s": pN / pw, "segments_per_s": pst["segments"] / pw, "ms_per_step": pw * 1e3,
                 "histories": pN}
        log(f"plain: {pw * 1e3:.1f} ms for {pN} histories")
        pctx.close()
    ctx.close()

    eigen = None
    if not args.no_eigen and args.workload == "c2":
        eigen = bench_eigen(args, tuning)

    assembly = None
    if not args.no_assembly and args.workload == "c2":
        assembly = bench_assembly(args)

    serving = None
    if not args.no_serve and args.workload == "c2":
        serving = bench_serving(args, tuning)

    c3 = None
    if not args.no_c3 and args.workload == "c2":
        c3 = bench_large(args, tuning, "c3")

    c5 = None
    if not args.no_c5 and args.workload == "c2":
        c5 = bench_large(args, tuning, "c5")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, N)
        log(f"cpu: {cpu['value']:.3e} histories/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": hist_s,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "description": desc, "n_elem": n_elem, "element": element,
                   "histories_per_step": N, "materials": n_mat, "nuclides": n_nuc,
                   "grid_points": int(prob.nuclides[0].energy.size),
                   "source": "volume Watt" + ("" if prob.source.angle == abi.ANGLE_ISOTROPIC else ", linear in mu"),
                   "boundary": "vacuum", "tally": "per-element track-length flux",
                   "mode": "batched", "precision": "fp64", "parallelism": "1 GPU", "tuning": tuning},
        "segments_per_s": seg_s,
        "segments_per_history": segs / (N * args.steps),
        "hex_segment_fraction": sum(s["hex_segments"] for s in steps) / max(1, segs),
        "collisions_per_history": colls / (N * args.steps),
        "plain_gpu": plain,
        "speedup_vs_plain": (hist_s / plain["histories_per_s"]) if plain else None,
        "mixed_precision": mixed,
        "mixed_speedup_vs_plain": (mixed["histories_per_s"] / plain["histories_per_s"]) if (plain and mixed) else None,
        "roofline": roof,
        "atomic_roofline": atomic_roof,
        "gather_roofline": gather_roof,
        "serving": serving,
        "eigenvalue": eigen,
        "beavrs_assembly": assembly,
        "c3": c3,
        "c5": c5,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
