"""Build the HIP library libbsa.so in-tree for gfx950 (hipcc cross-compiles
without a GPU).  Usage: python -m cucodeml_amd.build"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbsa.so")
SOURCES = ["bsa_engine.hip", "bsa_forest.hip", "bsa_kernels.hip", "bsa_noalloc.hip", "bsa_model.cpp",
           "bsa_fit.cpp",
           "bsa_batch.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # hipcc -O3 forms packed FP32 ops by itself; they are kept off on purpose
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-pthread",
    # host code (the model set-up's eigen systems) with AVX2 and FMA
    "-Xarch_host", "-mavx2", "-Xarch_host", "-mfma",
]


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libbsa.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
