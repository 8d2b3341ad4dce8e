"""Compile libhx.so (HIP kernels + C ABI) for gfx950, in tree.

Flags:
  -ffp-contract=off                       arithmetic that must match bit for bit
                                          is written with explicit fmaf only
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division/sqrt (Woop, Morton scale)
  -fno-gpu-flush-denormals-to-zero        IEEE denormals, as on the host
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                          no packed-FP32 instructions
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhx.so")
ARCH = os.environ.get("HX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hx_build.hip", "hx_ploc.hip", "hx_atr.hip", "hx_hlbvh.hip", "hx_insertion.hip", "hx_presplit.hip", "hx_batch.hip", "hx_sort.hip", "hx_trace.hip", "hx_diff.hip", "hx_noalloc.hip", "hx_raygen.hip", "hx_wide.hip", "hx_api.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in cand or os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose=False):
    """Compile every source and link libhx.so; raises on failure."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [_hipcc()] + HIP_FLAGS + lang + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
