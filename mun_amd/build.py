"""Build libmun_sem.so in-tree with hipcc for gfx950.

Usage: python -m mun_amd.build   (also called by __graft_entry__.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmun_sem.so")
SOURCES = [os.path.join(CSRC, "sem_lib.hip")]
HEADERS = [os.path.join(CSRC, "sem_kernels.h"), os.path.join(HERE, "..", "include", "sem_capi.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False, extra=()) -> str:
    """Compile every HIP source into mun_amd/libmun_sem.so (always rebuilds)."""
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           # packed FP32 instructions stay disabled: plain float code only
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
           "-o", LIB + ".tmp", *SOURCES, *extra]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
    print("built", LIB)
