"""Build the HIP library ``libcahip.so`` in-tree for gfx950.

Run as ``python -m cahip.build`` from ``cellularautomatacore.jl_amd/`` or call
:func:`build_library`.  Each translation unit under csrc/ compiles in its own
hipcc process (in parallel); the objects link into libcahip.so next to this
file, so that the library travels with the source tree.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libcahip.so")
UNITS = ["cahip.hip", "forest.hip", "lenia.hip", "life_r1.hip", "life_r2.hip", "life_r3.hip",
         "life_r4.hip", "life_r5.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every FP64/FP32 sum is written as it should round: no silent contraction
    # into fused multiply-adds (the kernels call fma explicitly where they want one)
    "-ffp-contract=off",
    # packed-FP32 instructions are left out of generated code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _compile(unit, verbose):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit}:\n" + res.stdout + res.stderr)
    return obj


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u))]
    with cf.ThreadPoolExecutor(max_workers=len(units)) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose), units))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
