"""Builds emdee_amd/libemdee.so for gfx950 (MI355X).

Kernels (csrc/*.hip) are compiled by hipcc for gfx950; the host driver
(csrc/*.cpp) is plain C++ built by g++ without FMA contraction, so that the
host-side arithmetic (model set-up, random momenta) is IEEE-exact and
reproducible.  hipcc links everything into one shared library.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libemdee.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[0])} {cmd[-1]}")
    return r


def build(verbose=False, jobs=None):
    """Compiles every source (in parallel, one process per file) and links."""
    os.makedirs(OBJ, exist_ok=True)
    cmds, objs = [], []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", "-o", o, src])
        objs.append(o)
    for src in sorted(glob.glob(os.path.join(CSRC, "*.cpp"))):
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", "-o", o, src])
        objs.append(o)
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-Wl,-Bsymbolic", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
