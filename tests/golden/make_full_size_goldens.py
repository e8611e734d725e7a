"""Write tests/golden/full_size.json: SHA-256 digests of oracle trajectories at
the benchmark sizes, for bit-exact GPU checks that the oracle cannot rerun in a
test's time budget.

The states come from the CPU oracle (oracle/caref.c, ArrayBackend engine) on the
documented synthetic inputs; a digest covers the column-major bytes of the
whole state.  Run:  python tests/golden/make_full_size_goldens.py [names...]
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cellularautomatacore.jl_amd"))

import numpy as np  # noqa: E402

from cahip import model as M  # noqa: E402
from cahip.backend import make_desc, make_params  # noqa: E402
from cahip.problem import CAProblem  # noqa: E402
from oracle import caref  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_size.json")
NT = int(os.environ.get("GOLDEN_THREADS", "6"))


def digest(u):
    return hashlib.sha256(np.asfortranarray(u).tobytes(order="F")).hexdigest()


def run(prob, saves):
    d, keep = make_desc(prob, M.Synchronous())
    params = make_params(prob, prob.p)
    u, t, out = prob.u0, prob.tspan[0], {}
    for ts in saves:
        u = caref.advance_halo(d, u, t, ts - t, params, NT)
        t = ts
        out[str(ts)] = {"sha256": digest(u), "sum": int(u.astype(np.int64).sum()),
                        "count2": int((u == 2).sum())}
        print(f"  t={ts} sum={out[str(ts)]['sum']}", flush=True)
    return out


def ltl():
    u0 = caref.fill_bernoulli(1, (8192, 8192), 0.5)
    prob = CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)), u0, 200,
                     space=M.Torus(8192, 8192), seed=1)
    return {"desc": "LtL Moore(5) B34..45/S34..58 Torus 8192^2, Bernoulli(0.5) seed 1 "
                    "(generator ctr=(cellid,0,0xFFFFFFFF,0))",
            "saves": run(prob, [50, 200])}


def forest():
    z = np.zeros((16384, 16384), np.int8, order="F")
    prob = CAProblem(M.ForestFire(), z, 500, {"p": 0.03, "f": 1e-4},
                     space=M.Grid(16384, 16384, boundary=M.Reflecting()), seed=7)
    return {"desc": "forest fire VN(1) Reflecting 16384^2 all-empty p=0.03 f=1e-4 seed 7",
            "saves": run(prob, [100, 500])}


def batch_sides(n=256, seed=2026):
    """The batched workload's documented lattice sides: round(64 * 32^u), u from
    numpy's default_rng(2026), square."""
    u = np.random.default_rng(seed).random(n)
    return [int(round(64 * 32 ** x)) for x in u]


def batched_forest():
    """256 forest fires, trajectory i (1-based) with seed i (prob.seed = 0 and
    seed_trajectory), sides batch_sides(), Reflecting, all empty, p = 0.03,
    f = 1e-4, 300 steps: the digest of every final state."""
    out = []
    for i, s in enumerate(batch_sides()):
        z = np.zeros((s, s), np.int8, order="F")
        prob = CAProblem(M.ForestFire(), z, 300, {"p": 0.03, "f": 1e-4},
                         space=M.Grid(s, s, boundary=M.Reflecting()), seed=i + 1)
        d, keep = make_desc(prob, M.Synchronous())
        u = caref.advance_halo(d, z, 0, 300, make_params(prob, prob.p), NT)
        out.append({"side": s, "seed": i + 1, "sha256": digest(u), "sum": int(u.astype(np.int64).sum()),
                    "count2": int((u == 2).sum())})
        if i % 32 == 0:
            print(f"  trajectory {i + 1}: side {s}", flush=True)
    return {"desc": "256 forest fires VN(1) Reflecting, sides round(64*32^u) (u ~ default_rng(2026)), "
                    "seeds 1..256, all empty, p=0.03 f=1e-4, 300 steps",
            "final": out}


JOBS = {"ltl_8192": ltl, "forest_16384": forest, "batched_forest_256": batched_forest}

if __name__ == "__main__":
    names = sys.argv[1:] or list(JOBS)
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for n in names:
        t = time.time()
        print(n, flush=True)
        data[n] = JOBS[n]()
        data[n]["oracle_seconds"] = round(time.time() - t, 1)
        json.dump(data, open(OUT, "w"), indent=1)
