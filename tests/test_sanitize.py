"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/Makefile builds, with g++ -fsanitize=address,undefined:
  - san_oracle: the CPU oracle and the library's host mesh derivation
    (mcw_mesh.cpp) driven through fixed-source and eigenvalue batches,
    general tallies, element temperatures, mixed TET4/HEX8 meshes with split
    interfaces, and their setup errors;
  - the input driver (HIT reader, mesh generators, Exodus reader, MOOSE-shaped
    objects), run on every input under tests/inputs and on Exodus files
    corrupted here: truncated at many lengths, header bytes overwritten at
    seeded positions, a coordinate array shorter than num_nodes, huge counts.
A run passes when it exits normally (0, or 1 with the driver's error message)
and prints no sanitizer report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
BUILD = os.path.join(SAN, "_build")
INPUTS = os.path.join(ROOT, "tests", "inputs")
LIB = os.path.join(ROOT, "macaw_amd", "lib", "libmacaw_amd.so")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        pytest.skip("the HIP library is not built (python -m macaw_amd.build)")
    r = subprocess.run(["make", "-s", "-j4", "-C", SAN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout
    return BUILD


def clean(r):
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode in (0, 1), (r.returncode, out[-4000:])
    return out


def driver(built, *args):
    r = subprocess.run([os.path.join(built, "macaw-amd-opt")] + list(args), stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, env=ENV, timeout=300)
    clean(r)
    return r


def test_oracle_and_host_mesh(built):
    r = subprocess.run([os.path.join(built, "san_oracle")], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, env=ENV, timeout=600)
    clean(r)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(INPUTS) if f.endswith(".i")))
def test_driver_inputs(built, name, tmp_path):
    r = driver(built, "-i", os.path.join(INPUTS, name), "--check-input")
    assert r.returncode == 0 and "input ok" in r.stdout, r.stderr
    r = driver(built, "-i", os.path.join(INPUTS, name), "--dump-problem", str(tmp_path / "p.dump"))
    assert r.returncode == 0, r.stderr


def corrupt_inputs(tmp_path, blobs):
    inp = open(os.path.join(INPUTS, "beavrs_assembly_3d.i")).read()
    for k, blob in enumerate(blobs):
        e = tmp_path / f"bad{k}.e"
        e.write_bytes(blob)
        i = tmp_path / f"bad{k}.i"
        i.write_text(inp.replace("'31enr_16_in.e'", f"'{e}'"))
        yield i


def test_exodus_truncated_and_overwritten(built, tmp_path):
    good = open(os.path.join(INPUTS, "31enr_16_in.e"), "rb").read()
    rng = np.random.default_rng(77)
    header = 4096
    blobs = [good[:n] for n in sorted(set(rng.integers(1, len(good), 24).tolist()) | {4, 8, 16, 100, 1000})]
    for _ in range(60):
        b = bytearray(good)
        for pos in rng.integers(4, header, rng.integers(1, 4)):
            b[pos] = int(rng.integers(0, 256))
        blobs.append(bytes(b))
    b = bytearray(good)
    b[3] = 5  # read as CDF-5: every count becomes 64 bits wide
    blobs.append(bytes(b))
    errors = 0
    for i in corrupt_inputs(tmp_path, blobs):
        r = driver(built, "-i", str(i), "--check-input")
        errors += r.returncode == 1
    assert errors >= 29  # every truncation is refused with a message


def test_exodus_short_coordinates_and_huge_counts(built, tmp_path):
    from scipy.io import netcdf_file
    blobs = []
    # coordx on a dimension shorter than num_nodes
    path = tmp_path / "short.e"
    f = netcdf_file(str(path), "w", version=2)
    f.createDimension("num_dim", 2)
    f.createDimension("num_nodes", 40)
    f.createDimension("three", 3)
    f.createDimension("num_el_blk", 1)
    for nm in ("coordx", "coordy"):
        v = f.createVariable(nm, "d", ("three",))
        v[:] = [0.0, 1.0, 2.0]
    f.close()
    blobs.append(path.read_bytes())
    # a dimension length far beyond the file, and one whose product overflows 64 bits
    for big in (0x7FFFFFFF, 0x40000000):
        path = tmp_path / f"big{big}.e"
        f = netcdf_file(str(path), "w", version=2)
        f.createDimension("num_dim", 2)
        f.createDimension("num_nodes", 3)
        f.createDimension("a", 5)
        f.createDimension("b", 7)
        for nm in ("coordx", "coordy"):
            v = f.createVariable(nm, "d", ("num_nodes",))
            v[:] = [0.0, 1.0, 2.0]
        v = f.createVariable("connect1", "i", ("a", "b"))
        v[:] = 1
        f.close()
        raw = bytearray(path.read_bytes())
        # dimension list: tag, count, then (name, length) per dimension; set
        # the lengths of "a" and "b"
        for name in (b"a\0\0\0", b"b\0\0\0"):
            at = raw.index(b"\0\0\0\x01" + name) + 8
            raw[at:at + 4] = big.to_bytes(4, "big")
        blobs.append(bytes(raw))
    for i in corrupt_inputs(tmp_path, blobs):
        r = driver(built, "-i", str(i), "--check-input")
        assert r.returncode == 1 and "*** ERROR ***" in r.stderr, r.stderr
