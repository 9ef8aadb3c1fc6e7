"""Autotuner of the batched (persistent) transport schedule.

Times candidate schedules on the caller's own problem, on the device, and
returns the fastest.  Knobs: workgroup size, `lanes_min` (how many of a
wavefront's 64 lanes must still be stepping segments before the wavefront
pauses to run its collisions and refills), the register/occupancy variant of
the kernel (`waves_per_eu`), and the bank chunk (histories resident per
launch).  Every candidate produces bit-identical histories, so tuning never
changes results, only speed.

    best, table = autotune(ctx, n_particles)
    ctx.set_options(**best)
"""
import itertools
import json
import os
import time

DEFAULT_SPACE = {
    "block_size": [256, 128],
    "lanes_min": [24, 32, 40, 48, 56],
    "waves_per_eu": [0, 5],
}


def autotune(ctx, n_particles, space=None, repeats=2, batch0=1000, log=None):
    """Time every combination in `space` (dict of option -> values) on
    `n_particles` histories per batch; returns (best option dict, rows)."""
    from . import abi
    space = dict(DEFAULT_SPACE if space is None else space)
    keys = sorted(space)
    rows = []
    saved = {k: getattr(ctx.opts, k) for k in keys + ["mode"]}
    ctx.set_options(mode=abi.MODE_BATCHED)
    b = batch0
    for combo in itertools.product(*(space[k] for k in keys)):
        opts = dict(zip(keys, combo))
        ctx.set_options(**opts)
        ctx.run_batch(b, n_particles)  # warm (first launch of a variant)
        b += 1
        times = []
        for _ in range(repeats):
            t = time.perf_counter()
            st = ctx.run_batch(b, n_particles)
            times.append(time.perf_counter() - t)
            b += 1
        rows.append(dict(opts, seconds=min(times), kernel_ms=st["kernel_ms"]))
        if log:
            log(f"autotune {opts}: {min(times) * 1e3:.2f} ms")
    best = min(rows, key=lambda r: r["seconds"])
    ctx.set_options(**saved)
    return {k: best[k] for k in keys}, rows


def cached(path, key, fn):
    """Memoise a tuning result in a JSON file under `key`."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    if key not in data:
        data[key] = fn()
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    return data[key]
