"""Compiles the HIP library for gfx950 (hipcc, in-tree)."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgenomix_gx.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("gx_build.hip", "gx_sizer.hip", "gx_batch.hip")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("gx_prims.h", "gx_kmer.h", "gx_sizer.h")] + [
    os.path.join(ROOT, "include", f) for f in ("genomix_gx.h", "genomix_sizer.h", "genomix_batch.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False):
    """Always recompiles (the build check compiles every time)."""
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", LIB + ".tmp"] + SOURCES
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
