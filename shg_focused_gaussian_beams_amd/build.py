"""Builds the HIP library (libshg.so) and the SHG and SHG_sweep programs in-tree.

hipcc compiles for gfx950 only.  Packed FP32 instructions are disabled
(-target-feature -packed-fp32-ops) for every HIP translation unit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libshg.so")
CLI = os.path.join(HERE, "SHG")
SWEEP = os.path.join(HERE, "SHG_sweep")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
SOURCES = ["ssfm_kernels.hip", "thermal_kernels.hip", "batch_kernels.hip", "adjoint_kernels.hip",
           "thermal_adjoint_kernels.hip", "shg_capi.cpp", "shg_batch.cpp", "shg_adjoint.cpp", "shg_noalloc.hip",
           "shg_thermal_batch.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, jobs=None):
    """Compiles every translation unit to an object (in parallel, into _build/), links
    libshg.so, then the two programs.  Compiles every time it is called."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        futs = [pool.submit(_run, [hipcc, *HIP_FLAGS, "-c", "-x", "hip", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs])
    os.replace(LIB + ".tmp", LIB)
    for exe, src in ((CLI, "shg_cli.cpp"), (SWEEP, "shg_sweep.cpp")):
        _run(["g++", "-O2", "-std=c++17", "-pthread", "-o", exe, os.path.join(CSRC, src), "-L" + HERE, "-lshg",
              "-Wl,-rpath,$ORIGIN", "-I" + os.path.join(HERE, "..", "include")])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
