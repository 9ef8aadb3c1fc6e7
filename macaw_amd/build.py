"""Build recipe of the HIP library (gfx950) and, for tests, the CPU oracle.

    python -m macaw_amd.build            # library + oracle
Compiles every source under macaw_amd/csrc with hipcc into
macaw_amd/lib/libmacaw_amd.so.  Flags that are part of the numerical
contract: -ffp-contract=off (no FMA contraction, so device arithmetic
matches the host's IEEE operations one for one) and no fast-math.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmacaw_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCW_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-parameter"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(LIBDIR, "obj", os.path.basename(src) + ".o")
        cmd = [HIPCC] + COMMON + ["-I", os.path.join(ROOT, "include")]
        if src.endswith(".hip"):
            cmd += DEVICE + ["-x", "hip"]
        cmd += ["-c", src, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(HERE, "host")
DRIVER = os.path.join(LIBDIR, "macaw-amd-opt")


def build_driver():
    """The MOOSE-shaped host layer and its input driver (host C++ only, linked
    against the library)."""
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-parameter", "-I", os.path.join(ROOT, "include")] + srcs
         + ["-o", tmp, "-L", LIBDIR, "-lmacaw_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main():
    lib = build_library(verbose="-v" in sys.argv)
    print("built", lib)
    print("built", build_driver())
    if "--no-oracle" not in sys.argv:
        build_oracle()
        print("built oracle")


if __name__ == "__main__":
    main()
