"""Build libspimrender.so (ray caster) and libspimlearn.so (learned reconstruction) for
gfx950 in-tree with hipcc.

    python -m spimagine_amd.build

The libraries land next to this file (spimagine_amd/lib*.so) so they travel with the
source tree; nothing is installed into site-packages.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspimrender.so")
LEARN_LIB = os.path.join(HERE, "libspimlearn.so")
LIBRARIES = {LIB: [os.path.join(CSRC, "spimrender.hip")],
             LEARN_LIB: [os.path.join(CSRC, "spimlearn.hip")]}

# -ffp-contract=off + correctly rounded fp32 div/sqrt: the kernels' float32 operation
# sequence is fixed (every FMA is an explicit fmaf).  Packed FP32 instructions are
# disabled (-packed-fp32-ops).
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    p = os.path.join(rocm, "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build_one(lib, sources, verbose=False):
    tmp = lib + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                     "-o", tmp] + sources
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % lib)
    os.replace(tmp, lib)
    return lib


def build_library(verbose=False):
    """Builds every library; returns the ray caster's path."""
    for lib, sources in LIBRARIES.items():
        build_one(lib, sources, verbose)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
