"""Builds libpyfof_hip.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m py_fof_amd.build  (or build() from Python).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpyfof_hip.so")
SOURCES = ["prims.hip", "fof.hip", "fft.hip", "bound.hip", "serve.hip", "noalloc.hip"]
ARCH = os.environ.get("PYFOF_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FoF predicate and the bookkeeping sums must round
# like the FMA-free reference build. Packed FP32 instructions are disabled.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    cand = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", CSRC, "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libpyfof_hip.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
