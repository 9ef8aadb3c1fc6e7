"""Autotuner of the batched (persistent) transport schedule.

Times candidate schedules on the caller's own problem, on the device, and
returns the fastest.  Knobs: workgroup size, `lanes_min` (how many of a
wavefront's 64 lanes must still be stepping segments before the wavefront
pauses to run its collisions and refills), the register/occupancy variant of
the kernel (`waves_per_eu`), then, with those fixed, the bank size
(histories resident per launch), the sort cadence (source sites ordered by
material and element, or not) and the flux tally's form (`tally_stage`:
staged in LDS and summed by a second kernel, or one device atomic per
segment; it only applies to some tallies, meshes and kernel variants).  The search is staged: a full grid over the
first group, then a grid over the second with the first group's best held.
Every candidate produces bit-identical histories, and tallies that differ
only in the order of FP64 additions, so tuning never changes results, only
speed.

    best, table = autotune(ctx, n_particles)
    ctx.set_options(**best)
"""
import itertools
import json
import os
import time

DEFAULT_SPACE = {
    "block_size": [256, 128],
    "lanes_min": [16, 24, 32, 40, 48, 56],
    "waves_per_eu": [0, 5],
}
BANK_SPACE = {
    "bank_capacity": [1 << 22, 1 << 23, 1 << 24],
    "sort_cadence": [0, 1],
    "tally_stage": [0, -1],
}


def _grid(ctx, n_particles, space, fixed, repeats, b, log):
    keys = sorted(space)
    rows = []
    for combo in itertools.product(*(space[k] for k in keys)):
        opts = dict(fixed, **dict(zip(keys, combo)))
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
    return rows, b


def autotune(ctx, n_particles, space=None, bank_space=None, repeats=2, batch0=1000, log=None):
    """Time every combination of `space` (dict of option -> values), then of
    `bank_space` with the first stage's best held, on `n_particles`
    histories per batch; returns (best option dict, rows)."""
    from . import abi
    space = dict(DEFAULT_SPACE if space is None else space)
    bank_space = dict(BANK_SPACE if bank_space is None else bank_space)
    keys = sorted(space) + sorted(bank_space)
    saved = {k: getattr(ctx.opts, k) for k in keys + ["mode"]}
    ctx.set_options(mode=abi.MODE_BATCHED)
    rows, b = _grid(ctx, n_particles, space, {}, repeats, batch0, log)
    best = {k: min(rows, key=lambda r: r["seconds"])[k] for k in sorted(space)}
    if bank_space:
        rows2, b = _grid(ctx, n_particles, bank_space, best, repeats, b, log)
        rows += rows2
        best = {k: min(rows2, key=lambda r: r["seconds"])[k] for k in keys}
    ctx.set_options(**saved)
    return best, rows


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
