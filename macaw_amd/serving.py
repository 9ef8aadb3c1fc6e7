"""Open- and closed-loop drivers of the request server (mcw_server_*).

`run_open_loop` submits each request at its arrival time (seconds after the
start) and polls the server between arrivals, as a front end feeding the
device would; `run_closed_loop` submits everything at once and measures the
saturated throughput.  Latency is the server's own clock: completion (the
flux has landed in host memory) minus submission.
"""
from __future__ import annotations

import time
from typing import Dict, List

import numpy as np


def _drain(server, out, keep_flux):
    n = server.poll()
    for _ in range(n):
        st, flux = server.result(copy=keep_flux)
        st["latency_ms"] = st["t_done_ms"] - st["t_submit_ms"]
        st["queue_ms"] = st["t_admit_ms"] - st["t_submit_ms"]
        if keep_flux:
            st["flux"] = flux
        out[st["request_id"]] = st


def run_open_loop(server, requests, keep_flux=False, log=None, progress_s=10.0) -> Dict[int, dict]:
    reqs = sorted(requests, key=lambda r: r.arrival_s)
    out: Dict[int, dict] = {}
    t0 = time.perf_counter()
    nxt = 0
    last = t0
    while len(out) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival_s <= now:
            r = reqs[nxt]
            server.submit(r.request_id, r.mesh_slot, r.n_particles, r.source)
            nxt += 1
        _drain(server, out, keep_flux)
        if log and time.perf_counter() - last > progress_s:
            last = time.perf_counter()
            log(f"serving: {nxt} submitted, {len(out)} done at t={last - t0:.1f}s")
        if nxt < len(reqs) and nxt == len(out):  # idle until the next arrival
            time.sleep(min(1e-3, max(0.0, reqs[nxt].arrival_s - (time.perf_counter() - t0))))
    return out


def run_closed_loop(server, requests, keep_flux=False) -> Dict[int, dict]:
    out: Dict[int, dict] = {}
    for r in requests:
        server.submit(r.request_id, r.mesh_slot, r.n_particles, r.source)
    while len(out) < len(requests):
        _drain(server, out, keep_flux)
    return out


def summarize(results: Dict[int, dict], wall_s: float) -> dict:
    lat = np.array([r["latency_ms"] for r in results.values()])
    hist = sum(r["histories"] for r in results.values())
    segs = sum(r["segments"] for r in results.values())
    return {"requests": len(results), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "histories": int(hist), "segments": int(segs), "wall_s": wall_s,
            "histories_per_s": hist / wall_s, "segments_per_s": segs / wall_s}
