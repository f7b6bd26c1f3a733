"""Builds libkpt_hip.so in-tree for gfx950 with hipcc.

IEEE build: -ffp-contract=off (no FMA contraction) and no fast-math on both the
host and device passes, so the kernels' float results match the CPU oracle;
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib", "libkpt_hip.so")

SOURCES = [
    os.path.join(CSRC, "hip", "kpt_kernels.hip"),
    os.path.join(CSRC, "hip", "kpt_wavefront.hip"),
    os.path.join(CSRC, "hip", "kpt_post.hip"),
    os.path.join(CSRC, "hip", "kpt_cache.hip"),
    os.path.join(CSRC, "hip", "kpt_abi.hip"),
    os.path.join(CSRC, "hip", "kpt_tlas.hip"),
    os.path.join(CSRC, "hip", "kpt_blas.hip"),
    os.path.join(CSRC, "host", "kpt_build.cpp"),
    os.path.join(CSRC, "host", "kpt_image.cpp"),
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES + ["-lpthread", "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libkpt_hip.so")
    os.replace(tmp, OUT)
    return OUT


INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def compile_cpp_app(sources, out, extra=()):
    """Compiles and links a host C++ application against include/ and the
    in-tree libkpt_hip.so (rpath set, so it runs without LD_LIBRARY_PATH)."""
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    libdir = os.path.dirname(OUT)
    cmd = (["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE] + list(sources) + list(extra) +
           ["-L", libdir, "-lkpt_hip", f"-Wl,-rpath,{libdir}", "-o", out])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"g++ failed building {out}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
