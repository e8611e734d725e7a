"""Build the HIP library ``libcahip.so`` in-tree for gfx950.

Run as ``python -m cahip.build`` from ``cellularautomatacore.jl_amd/`` or call
:func:`build_library`.  Each translation unit under csrc/ compiles in its own
hipcc process (in parallel); the objects link into libcahip.so next to this
file, so that the library travels with the source tree.  With
``incremental=True`` (``python -m cahip.build -i``) a unit is compiled again
only when it or a header it includes (the .d file hipcc wrote beside its
object) is newer than its object.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libcahip.so")
UNITS = ["cahip.hip", "forest.hip", "forest_batch.hip", "forest_gated.hip", "lenia.hip", "life_r1.hip",
         "life_r2.hip", "life_r3.hip", "life_r4.hip", "life_r5.hip", "life_r6.hip"]

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


def _up_to_date(obj, dep):
    """Whether obj is newer than every file its dependency list names."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    with open(dep) as f:
        text = f.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    t = os.path.getmtime(obj)
    return all(os.path.exists(x) and os.path.getmtime(x) <= t for x in files)


def _compile(unit, verbose, incremental=False):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
    dep = obj[:-2] + ".d"
    if incremental and _up_to_date(obj, dep):
        return obj
    cmd = [_hipcc(), *HIPCC_FLAGS, "-MMD", "-MF", dep, "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit}:\n" + res.stdout + res.stderr)
    return obj


def build_library(verbose=False, incremental=False):
    os.makedirs(OBJ, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u))]
    with cf.ThreadPoolExecutor(max_workers=len(units)) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose, incremental), units))
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
    print(build_library(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
