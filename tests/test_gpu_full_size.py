"""Whole trajectories at the benchmark sizes, bit-exact against digests of the
oracle's own runs (tests/golden/full_size.json, written by
tests/golden/make_full_size_goldens.py from the CPU oracle)."""
import hashlib

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend, check
from cahip.problem import CAProblem
from helpers import golden

pytestmark = pytest.mark.gpu
G = golden("full_size.json")


def digest(u):
    return hashlib.sha256(np.asfortranarray(u).tobytes(order="F")).hexdigest()


def run_and_check(be, prob, st, saves, want):
    t = prob.tspan[0]
    for ts in saves:
        be.advance(st, prob, M.Synchronous(), prob.p, t, ts - t)
        t = ts
        u = be.current_state(st)
        w = want[str(ts)]
        assert int(u.astype(np.int64).sum()) == w["sum"], f"t={ts}: cell sum differs"
        assert int((u == 2).sum()) == w["count2"], f"t={ts}: fire count differs"
        assert digest(u) == w["sha256"], f"t={ts}: state differs"


def test_ltl_8192_200_steps():
    be = HIPBackend()
    rule = M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5))
    prob = CAProblem(rule, np.zeros((8192, 8192), np.int8), 200, space=M.Torus(8192, 8192), seed=1)
    st = be.allocate_state(prob, M.Synchronous())
    try:
        # the documented generator, on the device: Bernoulli(0.5) from seed 1
        check(be.lib.cah_state_fill_bernoulli(st.handle, 1, 0.5))
        run_and_check(be, prob, st, [50, 200], G["ltl_8192"]["saves"])
    finally:
        be.free_state(st)


def test_forest_16384_500_steps():
    be = HIPBackend()
    z = np.zeros((16384, 16384), np.int8, order="F")
    prob = CAProblem(M.ForestFire(), z, 500, {"p": 0.03, "f": 1e-4},
                     space=M.Grid(16384, 16384, boundary=M.Reflecting()), seed=7)
    st = be.allocate_state(prob, M.Synchronous())
    try:
        run_and_check(be, prob, st, [100, 500], G["forest_16384"]["saves"])
    finally:
        be.free_state(st)
