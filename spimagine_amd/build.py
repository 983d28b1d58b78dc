"""Build libspimrender.so (ray caster) and libspimlearn.so (learned reconstruction) for
gfx950 in-tree with hipcc.

    python -m spimagine_amd.build

The libraries land next to this file (spimagine_amd/lib*.so) so they travel with the
source tree; nothing is installed into site-packages.  The ray caster is compiled as
several objects in parallel: the C ABI (spimrender.hip) and launch_typed.hip once per
voxel type and power-of-two rule (the kernel templates' instantiations), then linked.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspimrender.so")
LEARN_LIB = os.path.join(HERE, "libspimlearn.so")

# The ray-marching kernels (launch_typed.hip) run at two waves per SIMD (their LDS slabs
# bound occupancy), so the scheduler may spend registers on latency: LLVM's max-ILP
# strategy issues a sample's LDS reads further ahead of their use (1.5% on the 1024^2 max
# projection, 3% on the iso march; the instruction sequence per sample is unchanged).  The
# persistent batch launch (0.7% slower with it) and the plain gather kernels (the nearest
# one 23% slower) are compiled apart (SPR_PART 1) with the default scheduler.
SCHED_ILP = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]

# (object name, source, extra flags)
RENDER_OBJECTS = [("spimrender", "spimrender.hip", [])] + [
    ("launch_v%d_p%d" % (v, p2), "launch_typed.hip",
     ["-DSPR_VOXEL=%d" % v, "-DSPR_POW2=%d" % p2, "-DSPR_PART=0"] + SCHED_ILP)
    for v in (0, 1, 2) for p2 in (0, 1)] + [
    ("batch_v%d_p%d" % (v, p2), "launch_typed.hip",
     ["-DSPR_VOXEL=%d" % v, "-DSPR_POW2=%d" % p2, "-DSPR_PART=1"])
    for v in (0, 1, 2) for p2 in (0, 1)]
LEARN_OBJECTS = [("spimlearn", "spimlearn.hip", [])]
LIBRARIES = {LIB: RENDER_OBJECTS, LEARN_LIB: LEARN_OBJECTS}

# -ffp-contract=off + correctly rounded fp32 div/sqrt: the kernels' float32 operation
# sequence is fixed (every FMA is an explicit fmaf).  Packed FP32 instructions are
# disabled (-packed-fp32-ops).
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    p = os.path.join(rocm, "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % what)


def compile_object(name, source, flags, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, name + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + flags + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                                "-c", "-o", obj, os.path.join(CSRC, source)]
    _run(cmd, obj, verbose)
    return obj


def build_library(verbose=False, jobs=None):
    """Builds every library (objects compiled in parallel); returns the ray caster's path."""
    jobs = jobs or min(8, os.cpu_count() or 1)
    work = [(lib, o) for lib, objs in LIBRARIES.items() for o in objs]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {(lib, o[0]): ex.submit(compile_object, o[0], o[1], o[2], verbose) for lib, o in work}
        objs = {k: f.result() for k, f in futs.items()}
    for lib, specs in LIBRARIES.items():
        tmp = lib + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
              [objs[(lib, o[0])] for o in specs]
        _run(cmd, lib, verbose)
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
