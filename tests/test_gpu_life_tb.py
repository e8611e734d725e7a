"""The multi-step Moore(1) Life kernel (k_life_tb: up to 16 steps a launch, one
bit a cell in LDS) against the oracle at every saved state: step counts around
the 16-step launch, the three tile shapes (lattice sizes below 2^20, 2^24 and
above), sizes that are not multiples of 32, every edge kind including
Constant(2) (not alive) and mixed edges, Life rules with B0 and S8,
non-binary initial states, and batches of lattices (k_life_tb_batch)."""
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
    check(oracle, prob, [0, 1, 2, 9, 25, 41, 58, 60])


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


def check_batch(oracle, probs):
    t0, t1 = probs[0].tspan
    outs = HIPBackend().solve_batch(probs, M.Synchronous())
    for i, (p, out) in enumerate(zip(probs, outs)):
        d, keep = desc_for(p, M.Synchronous(), field_method=0)
        ref = oracle.advance_halo(d, p.u0, t0, t1 - t0, params_for(p), nthreads=NT)
        if not np.array_equal(out, ref):
            bad = np.argwhere(out != ref)
            raise AssertionError(f"member {i} {p.space.dims}: {len(bad)} cells differ, first {bad[0].tolist()}")


@pytest.mark.parametrize("steps", [1, 2, 16, 17, 35])
def test_batch_uneven_lattices(oracle, steps):
    """A batch of Moore(1) Life lattices of uneven sizes, edges and rules (the
    smallest tile shape: the batch has just over 1024 tiles of 64 x 16 cells)."""
    rng = np.random.default_rng(8 + steps)
    rules = ["conway", "B36/S23", "B2/S", "B3678/S34678"]
    dims_list = [(31, 17), (300, 211), (64, 64), (517, 60), (1000, 777), (96, 1030)]
    probs = []
    for i, dims in enumerate(dims_list):
        u = (rng.random(dims) < 0.4).astype(np.int8)
        probs.append(CAProblem(M.Life(rules[i % 4]), u, (3, 3 + steps),
                               space=M.Lattice(dims, EDGES[i % len(EDGES)])))
    check_batch(oracle, probs)


@pytest.mark.parametrize("side,n", [(2048, 4), (6144, 2)])
def test_batch_larger_tiles(oracle, side, n):
    """Batches whose tiles are 128 x 32 (4 lattices of 2048^2) and 512 x 128
    (2 lattices of 6144^2) cells."""
    rng = np.random.default_rng(side)
    probs = []
    for i in range(n):
        u = (rng.random((side, side)) < 0.35).astype(np.int8)
        probs.append(CAProblem(M.Life("conway"), u, 17,
                               space=M.Lattice((side, side), EDGES[i % len(EDGES)])))
    check_batch(oracle, probs)
