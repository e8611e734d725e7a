"""Benchmark of the synchronous sweep on one MI355X.

Workload (one "step"): one synchronous step of each of
  - Larger-than-Life Moore(5) B34..45/S34..58, Torus 8192 x 8192, Int8,
    Bernoulli(0.5) initial state from seed 1;
  - the forest fire, VonNeumann(1), Reflecting 16384 x 16384, Int8, all empty,
    p = 0.03, f = 1e-4, Philox seed 7;
both lattices resident in HBM.  `value` = cell updates per second over both, in
GCUPS.  Kernel durations come from HIP events recorded on each state's own
stream around its launches.  The CPU baseline is the oracle's restatement of
the reference's threaded ArrayBackend, timed on a few steps of each lattice.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cellularautomatacore.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def make_states(be):
    from cahip import model as M
    from cahip.backend import check
    from cahip.problem import CAProblem
    ltl = CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)),
                    np.zeros((8192, 8192), np.int8), 200, space=M.Torus(8192, 8192), seed=1)
    forest = CAProblem(M.ForestFire(), np.zeros((16384, 16384), np.int8), 500,
                       {"p": 0.03, "f": 1e-4}, space=M.Grid(16384, 16384, boundary=M.Reflecting()),
                       seed=7)
    s_ltl = be.allocate_state(ltl, M.Synchronous())
    check(be.lib.cah_state_fill_bernoulli(s_ltl.handle, 1, 0.5))
    s_for = be.allocate_state(forest, M.Synchronous())
    return (ltl, s_ltl), (forest, s_for)


def cpu_baseline(nthreads, steps=2):
    """The oracle's ArrayBackend restatement on the host cores, a few steps of each lattice."""
    from cahip import model as M
    from cahip.backend import make_desc, make_params
    from cahip.problem import CAProblem
    from oracle import caref
    caref.load()
    u = caref.fill_bernoulli(1, (8192, 8192), 0.5)
    ltl = CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)), u, 1,
                    space=M.Torus(8192, 8192), seed=1)
    d1, k1 = make_desc(ltl, M.Synchronous())
    t = time.perf_counter()
    caref.advance_halo(d1, u, 0, steps, None, nthreads)
    t_ltl = (time.perf_counter() - t) / steps
    del u, ltl
    z = np.zeros((16384, 16384), np.int8, order="F")
    fo = CAProblem(M.ForestFire(), z, 1, {"p": 0.03, "f": 1e-4},
                   space=M.Grid(16384, 16384, boundary=M.Reflecting()), seed=7)
    d2, k2 = make_desc(fo, M.Synchronous())
    t = time.perf_counter()
    caref.advance_halo(d2, z, 0, steps, make_params(fo, fo.p), nthreads)
    t_for = (time.perf_counter() - t) / steps
    cells = 8192 * 8192 + 16384 * 16384
    return {"value": cells / (t_ltl + t_for) / 1e9, "unit": "GCUPS", "cores": nthreads,
            "kind": "port",
            "sample": f"{steps} steps of each lattice (8192^2 LtL, 16384^2 forest fire) with the "
                      "oracle's C restatement of ArrayBackend(threaded=true): halo-padded, serial "
                      "halo fill, min(n_last, 4*threads) slabs",
            "ltl_ms_per_step": t_ltl * 1e3, "forest_ms_per_step": t_for * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the sweep on a single MI355X "
              "(the multi-GPU slab decomposition is design only)", file=sys.stderr)
        sys.exit(2)

    from cahip.backend import HIPBackend, check, make_params
    be = HIPBackend(0)
    (ltl, s_ltl), (forest, s_for) = make_states(be)
    ev = HipEvents()
    from cahip import model as M
    sync = M.Synchronous()
    pf = make_params(forest, forest.p)
    pl = make_params(ltl, None)

    def run(st, params, t0, n):
        check(be.lib.cah_advance(st.handle, t0, n, ctypes.byref(params)))

    run(s_ltl, pl, 0, args.warmup)
    run(s_for, pf, 0, args.warmup)
    ev.sync()

    K = args.steps
    # each lattice timed on its own (the other idle), with HIP events recorded on
    # the state's own stream around its K launches
    str_ltl, str_for = be.stream(s_ltl), be.stream(s_for)
    a1, b1, a2, b2 = ev.event(), ev.event(), ev.event(), ev.event()
    t_wall = time.perf_counter()
    ev.record(a1, str_ltl)
    run(s_ltl, pl, args.warmup, K)
    ev.record(b1, str_ltl)
    ev.sync()
    ev.record(a2, str_for)
    run(s_for, pf, args.warmup, K)
    ev.record(b2, str_for)
    ev.sync()
    wall = time.perf_counter() - t_wall
    ms_ltl = ev.elapsed_ms(a1, b1) / K
    ms_for = ev.elapsed_ms(a2, b2) / K

    c_ltl, c_for = 8192 * 8192, 16384 * 16384
    ms_step = ms_ltl + ms_for
    value = (c_ltl + c_for) / (ms_step * 1e-3) / 1e9
    # dominant kernel: the forest fire sweep (one launch per step)
    bytes_per_launch = 2 * c_for
    achieved = bytes_per_launch / (ms_for * 1e-3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "forest_vn1_16384_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": "GCUPS + % HBM roofline (synchronous sweep)",
        "value": round(value, 3), "unit": "GCUPS", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round(wall * 1e3 / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8", "data": "synthetic",
        "config": {"workload": "LtL Moore(5) B34..45/S34..58 Torus 8192^2 Bernoulli(0.5) seed 1 "
                               "+ forest fire VN(1) Reflecting 16384^2 all-empty p=0.03 f=1e-4 seed 7",
                   "global_batch": 2, "seq_len": None, "parallelism": "none (1 GPU)"},
        "workloads": {
            "ltl_8192": {"gcups": round(c_ltl / ms_ltl / 1e6, 3), "ms_per_step": round(ms_ltl, 5),
                         "hbm_gbs_algorithmic": round(2 * c_ltl / ms_ltl / 1e6, 1)},
            "forest_16384": {"gcups": round(c_for / ms_for / 1e6, 3),
                             "ms_per_step": round(ms_for, 5),
                             "hbm_gbs_algorithmic": round(2 * c_for / ms_for / 1e6, 1)},
        },
        "roofline": {"bound": "hbm", "kernel": "k_forest_vn1", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
    }
    be.free_state(s_ltl)
    be.free_state(s_for)
    if not args.no_cpu:
        n = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
