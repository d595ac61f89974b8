"""Builds libksvd_amd.so in-tree with hipcc for gfx950 (no cmake needed).

    python -m ksvd_amd.build          # compile every source, link the library
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libksvd_amd.so")
SOURCES = ["omp_qform_t64.hip", "omp_qform_t48.hip", "omp_qform_t32.hip", "omp_qform_t16.hip", "omp_qform_t8.hip",
           "gemm.hip", "omp.hip", "omp_qform.hip", "dict_update.hip", "adjoint.hip", "capi.cpp",
           "matio.cpp"]
MAIN = os.path.join(HERE, "ksvd_main")  # CLI driver (ksvd_main.cpp) linked against the library

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KSVD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off: the engine relies on scalar FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Per-source extra flags.  The register-basis Batch-OMP kernel (omp_qform.hpp,
# one translation unit per basis size, the longest first) unrolls all of
# its (up to 64) steps so that every index into its register-held basis is a
# compile-time constant; LLVM's default size limit for `#pragma unroll` would
# leave the loop rolled and push the basis to scratch memory.
UNROLL = ["-mllvm", "-pragma-unroll-threshold=500000"]
EXTRA = {f"omp_qform_t{tt}.hip": UNROLL for tt in (8, 16, 32, 48, 64)}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = MAIN + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "ksvd_main.cpp"), "-o", tmp,
           f"-L{HERE}",
           "-lksvd_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ksvd_main build failed:\n{r.stderr}")
    os.replace(tmp, MAIN)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
