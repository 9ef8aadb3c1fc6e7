"""bench.py's host-side arithmetic, on the CPU: the algorithmic byte model,
the roofline of a launch with and without the staged tally's sums, the
traffic figure read from the committed counter profile, and the one-GPU
guard."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def step(segments, collisions, kernel_ms, stage_ms=0.0, hex_segments=0, launches=1):
    return {"segments": segments, "collisions": collisions, "kernel_ms": kernel_ms, "stage_ms": stage_ms,
            "hex_segments": hex_segments, "kernel_launches": launches}


def test_algorithmic_bytes_terms():
    # 92 B per history, B_seg per segment, 64 n + 64 per collision
    b = bench.algorithmic_bytes({"histories": 10, "segments": 100, "collisions": 3}, 4, 276.0)
    assert b == 92 * 10 + 276 * 100 + (64 * 4 + 64) * 3


def test_roofline_of_unstaged_and_staged():
    n = 1000
    plain = bench.roofline_of([step(50_000, 900, 2.0)], n, 4, 276.0)
    staged = bench.roofline_of([step(50_000, 900, 2.0, stage_ms=0.25)], n, 4, 276.0)
    want = bench.algorithmic_bytes({"histories": n, "segments": 50_000, "collisions": 900}, 4, 276.0)
    for r in (plain, staged):
        assert r["alg_bytes_per_launch"] == pytest.approx(want)
        assert r["launch_ms"] == pytest.approx(2.0)
        assert r["achieved"] == pytest.approx(want / 2e-3 / 1e9)
        assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS)
    # the staged launch's time is split into the transport kernel and the sums
    assert "stage_ms" not in plain and "transport_ms" not in plain
    assert staged["stage_ms"] == pytest.approx(0.25)
    assert staged["transport_ms"] == pytest.approx(1.75)


def test_roofline_of_counts_hex_segments_and_launches():
    r = bench.roofline_of([step(1000, 0, 4.0, hex_segments=400, launches=2)], 10, 4, 276.0, b_hex=348.0)
    want = bench.algorithmic_bytes({"histories": 10, "segments": 600, "collisions": 0}, 4, 276.0) + 348.0 * 400
    assert r["alg_bytes_per_launch"] == pytest.approx(want / 2)
    assert r["launch_ms"] == pytest.approx(2.0)
    assert r["launches_per_step"] == 2


def test_read_traffic_is_the_staged_launch():
    t = bench.read_traffic("c2_tet4_1008k_vacuum")
    text = open(os.path.join(ROOT, "profiles", "k_persistent_c2_tet4_1008k_vacuum_pmc.txt")).read()
    first = [l for l in text.splitlines() if l.startswith("traffic_bytes_per_launch")]
    assert len(first) == 1 and t == float(first[0].split("=")[1].split()[0])
    assert 2.5e11 < t < 3.5e11
    assert bench.read_traffic("no_such_workload") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
