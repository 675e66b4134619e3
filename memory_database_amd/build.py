"""Build libolapcore.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m memory_database_amd.build

The library lands next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "libolapcore.so")
DRIVER_PATH = os.path.join(HERE, "olapcore_test")
HARNESS_SRC = os.path.join(ROOT, "tests", "native", "compat_harness.cpp")
HARNESS_PATH = os.path.join(ROOT, "tests", "native", "compat_harness")

ARCH = os.environ.get("OLAP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["olap_kernels.hip", "olap_stage.hip", "olap_batch.hip", "olap_noalloc.hip",
           "olap_join.hip", "olap_predicate.hip", "olap_snowflake.hip",
           "olap_capi.cpp",
           "olapcore_compat.cpp",
           "olap_server.cpp"]

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep packed-FP32 instructions out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose: bool = False) -> str:
    """Compile every source (in parallel) and link one shared library; returns
    its path."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = _hipcc()
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as pool:
        futs = [pool.submit(_run, [hipcc, *HIP_FLAGS, *inc, "-c", "-o", o, s], verbose)
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB_PATH)
    drv = os.path.join(CSRC, "olapcore_driver.cpp")
    if os.path.exists(drv):
        cmd = [hipcc, "-O2", "-std=c++17", *inc, "-o", DRIVER_PATH, drv,
               f"-L{HERE}", "-lolapcore", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"]
        _run(cmd, verbose)
    # the GPU tests' caller of the reference-signature layer, linked the way a
    # caller of the reference's operators links the library
    if os.path.exists(HARNESS_SRC):
        cmd = [hipcc, "-O2", "-std=c++17", *inc, "-o", HARNESS_PATH, HARNESS_SRC,
               f"-L{HERE}", "-lolapcore", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../memory_database_amd"]
        _run(cmd, verbose)
    return LIB_PATH


if __name__ == "__main__":
    build_library(verbose="-q" not in sys.argv)
    print(LIB_PATH)
