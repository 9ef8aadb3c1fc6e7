#!/usr/bin/env python3
"""Benchmark: histories/s and segments/s of the particle-tracking hot path.

Workload (one "step" = one batch): unit-cube TET4 mesh from a Delaunay
triangulation of a jittered 54^3 lattice (1,008,162 tets), vacuum
boundaries, 4 Voronoi materials of 4 synthetic nuclides (50k-point grids),
uniform volumetric Watt source, 1e7 histories per batch, FP64, a per-element
track-length flux tally.  Inputs are synthetic and resident in HBM before
the timed region.

Reported beside the batched (persistent, fused) path: the plain event-based
GPU path on the same batch, the dominant kernel's roofline (algorithmic bytes
per launch over its HIP-event duration) and a CPU baseline (the FP64 oracle,
std::thread on the host cores, on a bounded sample of the same histories).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--histories N]
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
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
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
                "bank_gib": args.serve_bank_gib, "particles": "log-uniform [1e4, 1e7]",
                "sources": "point or box (1:1), Watt", "tally": "per-element track-length flux per request",
                "arrivals": "Poisson, seed 505", "precision": "fp64"})
    log(f"serving: p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, "
        f"{out['histories_per_s']:.3e} histories/s over {wall:.1f}s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=int, default=10_000_000)
    ap.add_argument("--lattice", type=int, default=54)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lanes-min", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving mix")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-calibrate", type=int, default=100, help="requests in the saturated capacity run")
    ap.add_argument("--serve-load", type=float, default=0.7, help="offered load as a fraction of capacity")
    ap.add_argument("--serve-inflight", type=int, default=32)
    ap.add_argument("--serve-bank-gib", type=float, default=4.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="N",
                    help="only run one streaming probe kernel over N doubles (reads 16N bytes, writes 48N "
                         "bytes): the known-size reference for rocprofv3 FETCH_SIZE/WRITE_SIZE scaling")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(histories shard by id range across GPUs with one tally all-reduce per batch; "
              f"that path is not built)", file=sys.stderr)
        sys.exit(2)

    from macaw_amd import abi, problems as P
    from macaw_amd.lib import Context

    if args.calibrate_copy:
        import numpy as np
        n = args.calibrate_copy
        c = Context()
        c.math_probe(np.linspace(0.1, 0.9, n), np.arange(n, dtype=np.uint64))
        print(json.dumps({"calibration": "k_math_probe", "n": n, "read_bytes": 16 * n, "write_bytes": 48 * n}))
        return

    t0 = time.perf_counter()
    log(f"generating workload: Delaunay lattice {args.lattice}^3")
    prob = P.c2_tet_vacuum(n_lattice=args.lattice, n_particles=args.histories)
    n_elem = prob.mesh.n_elem
    workload = f"c2_tet4_{n_elem // 1000}k_vacuum"
    log(f"mesh {n_elem} tets, {prob.mesh.n_nodes} nodes ({time.perf_counter() - t0:.1f}s)")

    ctx = Context(prob, mode=abi.MODE_BATCHED, block_size=args.block, lanes_min=args.lanes_min)
    N = args.histories
    tuning = {"block_size": ctx.opts.block_size, "lanes_min": ctx.opts.lanes_min,
              "waves_per_eu": ctx.opts.waves_per_eu}
    if not args.no_autotune:
        from macaw_amd.autotune import autotune
        t = time.perf_counter()
        tuning, rows = autotune(ctx, N, log=log)
        ctx.set_options(**tuning)
        log(f"autotune: best {tuning} ({time.perf_counter() - t:.1f}s)")
    for w in range(args.warmup):
        st = ctx.run_batch(100 + w, N)
        log(f"warmup {w}: {st['time_ms']:.1f} ms, {st['segments']} segments")
    steps = []
    t = time.perf_counter()
    for k in range(args.steps):
        steps.append(ctx.run_batch(k + 1, N))
    wall = time.perf_counter() - t
    for k, st in enumerate(steps):
        log(f"step {k}: {st['time_ms']:.1f} ms device, {st['segments']} segments, {st['collisions']} collisions, "
            f"leaks {st['leaks']}, lost {st['lost']}")
    segs = sum(s["segments"] for s in steps)
    colls = sum(s["collisions"] for s in steps)
    kern_ms = sum(s["kernel_ms"] for s in steps)
    launches = sum(s["kernel_launches"] for s in steps)
    hist_s = N * args.steps / wall
    seg_s = segs / wall

    # dominant kernel: k_persistent, one launch per batch
    agg = {"histories": N * args.steps, "segments": segs, "collisions": colls}
    b_alg = algorithmic_bytes(agg, 4)
    per_launch_bytes = b_alg / launches
    per_launch_s = kern_ms / launches / 1e3
    achieved = per_launch_bytes / per_launch_s / 1e9
    traffic = read_traffic(workload)

    plain = None
    if not args.no_plain:
        pctx = Context(prob, mode=abi.MODE_PLAIN)
        pctx.run_batch(200, N)  # warm
        t = time.perf_counter()
        pst = pctx.run_batch(1, N)
        pw = time.perf_counter() - t
        plain = {"histories_per_s": N / pw, "segments_per_s": pst["segments"] / pw, "ms_per_step": pw * 1e3}
        if pst["segments"] != steps[0]["segments"]:
            log("WARNING: plain and batched segment counts differ")
        log(f"plain: {pw * 1e3:.1f} ms per batch")
        pctx.close()
    ctx.close()

    serving = None
    if not args.no_serve:
        serving = bench_serving(args, tuning)

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
        "config": {"workload": workload, "n_elem": n_elem, "element": "TET4", "histories_per_step": N,
                   "materials": 4, "nuclides": 4, "grid_points": 50000, "source": "volume Watt",
                   "boundary": "vacuum", "tally": "per-element track-length flux", "mode": "batched",
                   "precision": "fp64", "parallelism": "1 GPU", "tuning": tuning},
        "segments_per_s": seg_s,
        "segments_per_history": segs / (N * args.steps),
        "collisions_per_history": colls / (N * args.steps),
        "plain_gpu": plain,
        "speedup_vs_plain": (hist_s / plain["histories_per_s"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_persistent<4>", "alg_bytes_per_launch": per_launch_bytes,
                     "launch_ms": per_launch_s * 1e3},
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
