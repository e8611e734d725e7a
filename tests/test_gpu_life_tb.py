"""The multi-step Moore(1) Life kernel (k_life_tb: up to 8 steps a launch, one
bit a cell in LDS) against the oracle at every saved state: step counts around
the 8-step launch, the three tile shapes (lattice sizes below 2^20, 2^24 and
above), sizes that are not multiples of 32, every edge kind including
Constant(2) (not alive) and mixed edges, Life rules with B0 and S8, and
non-binary initial states."""
import os

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, solve
from helpers import desc_for, params_for

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def check(oracle, prob, saveat):
    sol = solve(prob, M.Synchronous(), backend=HIPBackend(), saveat=saveat)
    d, keep = desc_for(prob, M.Synchronous(), field_method=0)
    u, t = prob.u0, sol.t[0]
    for k in range(1, len(sol.t)):
        u = oracle.advance_halo(d, u, t, sol.t[k] - t, params_for(prob), nthreads=NT)
        t = sol.t[k]
        if not np.array_equal(sol.u[k], u):
            bad = np.argwhere(sol.u[k] != u)
            raise AssertionError(f"t={t}: {len(bad)} cells differ, first {bad[0].tolist()}")


EDGES = [(M.Periodic(), M.Periodic()), (M.Reflecting(), M.Reflecting()), (M.Constant(0), M.Constant(1)),
         (M.Constant(1), M.Periodic()), (M.Reflecting(), M.Constant(2)), (M.Periodic(), M.Reflecting())]


@pytest.mark.parametrize("e", range(len(EDGES)))
@pytest.mark.parametrize("dims", [(31, 17), (517, 60), (1000, 777), (33, 1025)])
def test_conway_edges_and_sizes(oracle, e, dims):
    u = (np.random.default_rng(e * 7 + dims[0]).random(dims) < 0.4).astype(np.int8)
    prob = CAProblem(M.Life("conway"), u, 60, space=M.Lattice(dims, EDGES[e]))
    check(oracle, prob, [0, 1, 2, 9, 17, 25, 60])


@pytest.mark.parametrize("rule", ["B36/S23", "B2/S", "B3678/S34678", "B0/S8", "B0123478/S01234678"])
def test_life_rules(oracle, rule):
    dims = (300, 211)
    u = (np.random.default_rng(3).random(dims) < 0.3).astype(np.int8)
    prob = CAProblem(M.Life(rule), u, 20, space=M.Lattice(dims, (M.Reflecting(), M.Constant(1))))
    check(oracle, prob, [0, 7, 8, 20])


def test_non_binary_states(oracle, backend):
    # set_state! may bring values a Life rule never makes; only 1 counts as alive
    dims = (259, 130)
    u = np.random.default_rng(4).choice(np.array([0, 1, 2, -1, 3, 1], np.int8), dims)
    prob = CAProblem(M.Life("conway"), np.zeros(dims, np.int8), 11, space=M.Torus(*dims))
    st = backend.allocate_state(prob, M.Synchronous())
    try:
        backend.set_state(st, u)
        backend.advance(st, prob, M.Synchronous(), None, 0, 11)
        d, keep = desc_for(prob, M.Synchronous(), field_method=0)
        assert np.array_equal(backend.current_state(st), oracle.advance_plain(d, u, 0, 11))
    finally:
        backend.free_state(st)


@pytest.mark.parametrize("dims", [(2049, 1030), (4200, 4100)])
def test_larger_tiles(oracle, dims):
    u = (np.random.default_rng(5).random(dims) < 0.35).astype(np.int8)
    prob = CAProblem(M.Life("conway"), u, 18, space=M.Lattice(dims, (M.Periodic(), M.Constant(0))))
    check(oracle, prob, [0, 18])


def test_conway_512_long_run(oracle):
    u = oracle.fill_bernoulli(1, (512, 512), 0.5)
    prob = CAProblem(M.Life("conway"), u, 1000, space=M.Torus(512, 512))
    check(oracle, prob, [0, 333, 1000])
