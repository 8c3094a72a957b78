"""Builds pump_amd/libpump_amd.so (all HIP sources, gfx950) in-tree.

    python pump_amd/build.py

(Run it by path: importing the pump_amd package loads the built library,
which the build replaces.)

Every source is compiled with hipcc for gfx950 with FP contraction off (the
kernels reproduce the reference's uncontracted float/double arithmetic) and
packed-FP32 instruction selection disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpump_amd.so")

SOURCES = [
    "bvp.hip",
    "roadmap.hip",
    "edges.hip",
    "hsmc.hip",
    "batch.hip",
    "hsmc_bins.hip",
    "hsmc_mfma.hip",
    "mccp.hip",
    "host_gen.hip",
    "host_pipeline.hip",
    "search_device.hip",
    "smooth.hip",
    "problem_io.hip",
    "gmt.hip",
    "runtime.hip",
    "noalloc.hip",
    "noise.hip",
]

ARCH = os.environ.get("PUMP_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), "-c", *CFLAGS, os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    """Compile every source and link the shared library (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
