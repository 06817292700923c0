"""Build liblhl_hip.so (HIP, gfx950) in-tree with hipcc.

Usage: python lhl.jl_amd/build_lib.py   (or __graft_entry__.build()).
The shared library lands in lhl.jl_amd/lib/liblhl_hip.so; objects in
lhl.jl_amd/lib/obj/.  No GPU is needed: hipcc cross-compiles for gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "liblhl_hip.so")
SOURCES = ["lhl_reduce.hip", "lhl_blocked.hip", "lhl_shift.hip", "lhl_solve.hip", "lhl_batch.hip", "lhl_gemm.hip",
           "lhl_cplx.hip", "lhl_abi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32 (they have produced wrong
# results on these boxes).  Correctly rounded FP32 division/sqrt so the shift
# LU's divisions are IEEE in both precisions.
CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(LIBDIR, "obj", os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(7, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    # rocBLAS: the refinement residual's plain GEMM (J x) in one precision
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
