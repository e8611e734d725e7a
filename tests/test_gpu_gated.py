"""The gated band walks against the CPU oracle, bit for bit at every save: Life
and Generations (k_life_gated) and the forest fire (k_forest_gated) under the
Stochastic (Float64 and Float32 p) and Checkerboard gates and masks
(src/backends.jl:444-455, src/schemes.jl:74-93, src/rules.jl:446-456).

The shapes take both forms of the walk: sides that are multiples of the lane
width and at least 64 lanes wide (buffer loads with the boundary folded in, the
mask read a lane's group at a time) and the others (the general boundary map,
the mask byte by byte); short row chunks, so that half the waves walk up their
chunk and read the mask rows mirrored; every edge kind; launch shapes set with
cah_state_tune.  Cells outside a rule's states (which the solve API refuses)
go through the state API, and the oracle's plain sweep takes them as they are.
"""
import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, solve
from helpers import desc_for, ltl_rule, params_for, rng_state
from test_gpu_parity import assert_same, oracle_saves

pytestmark = pytest.mark.gpu

P, R, C = M.Periodic(), M.Reflecting(), M.Constant

F32 = np.float32


def mask_of(dims, seed, frac=0.2):
    return np.random.default_rng(seed).random(dims) >= frac


LIFE_CASES = [
    # rule, dims, edges, values, scheme, masked, tune
    ("brians_brain", (1024, 300), (P, P), [0, 1, 2], M.Synchronous(), False, None),
    ("brians_brain", (1000, 37), (R, C(2)), [0, 1, 2], M.Checkerboard(), True, None),
    ("star_wars", (2048, 64), (P, R), [0, 1, 2, 3], M.Stochastic(0.5), True, None),
    ("star_wars", (517, 61), (C(3), P), [0, 1, 2, 3], M.Stochastic(F32(0.3)), False, None),
    ("B23/S34/C5 Moore(2)", (768, 90), (R, R), [0, 1, 2, 3, 4], M.Checkerboard(), True, None),
    ("conway", (512, 512), (P, P), [0, 1], M.Stochastic(F32(0.3)), False, None),
    ("conway", (512, 200), (C(1), R), [0, 1], M.Synchronous(), True, (2, 7)),
    ("B36/S23 Moore(3)", (2050, 40), (C(1), C(0)), [0, 1], M.Stochastic(0.7), True, None),
    ("ltl", (1100, 97), (P, P), [0, 1], M.Checkerboard(), True, None),
    ("ltl", (1024, 80), (R, C(1)), [0, 1], M.Stochastic(0.45), False, None),
    ("brians_brain", (256, 256), (P, P), [0, 1, 2], M.Stochastic(1.0), True, (1, 5)),
    # states of 0 and 1 only at the start (the first sweeps count cells as they
    # are, and the flag clears once a Generations sweep has written 2)
    ("brians_brain", (640, 100), (C(0), C(1)), [0, 1], M.Stochastic(0.8), False, None),
    ("star_wars", (1024, 64), (P, P), [0, 1], M.Synchronous(), True, None),
    ("star_wars", (1024, 64), (R, P), [0, 1], M.Checkerboard(), False, (2, 9)),
    # radius 6 (counts past 127 in the byte lanes); radius 7 takes the window-sum kernel
    ("ltl6", (1100, 70), (P, R), [0, 0, 0, 1], M.Stochastic(0.6), True, None),
    ("B45/S345/C5 Moore(6)", (600, 90), (C(1), P), [0] * 40 + [1, 2, 3, 4], M.Checkerboard(), False, None),
    ("B45/S345/C5 Moore(6)", (517, 61), (R, C(2)), [0] * 40 + [1, 2, 3, 4], M.Stochastic(0.5), True, None),
    ("ltl7", (517, 61), (R, R), [0, 0, 0, 0, 1], M.Synchronous(), True, None),
    # Generations of at most 4 states on lattices of cells below 4 (the small-cell
    # rule from the second sweep on, or from the first where the set state is
    # scanned), with gates, masks, Constant(2) and Constant(3) edges and odd
    # shapes; a Constant(4) edge keeps the exact byte tests
    ("brians_brain", (1000, 37), (R, C(2)), [0, 1], M.Checkerboard(), True, None),
    ("brians_brain", (517, 61), (R, P), [0, 1, 2], M.Synchronous(), False, None),
    ("brians_brain", (512, 64), (C(3), P), [0, 1], M.Stochastic(0.5), True, None),
    ("star_wars", (600, 50), (C(4), R), [0, 1, 2, 3], M.Checkerboard(), False, None),
    ("star_wars", (1030, 70), (P, C(3)), [0, 1, 2, 3], M.Synchronous(), False, None),
    ("B3/S23/C4 Moore(2)", (512, 80), (R, C(2)), [0, 1, 2, 3], M.Stochastic(0.5), True, None),
    ("B3/S23/C2 Moore(2)", (768, 90), (C(1), P), [0, 1], M.Stochastic(0.5), True, None),
    ("B2/S/C3 Moore(3)", (1024, 70), (P, P), [0, 1, 2], M.Stochastic(F32(0.4)), True, None),
    # radii 3 and more without a mask take the walk that keeps no mask ring
    ("B36/S23 Moore(3)", (1030, 50), (P, C(0)), [0, 1], M.Checkerboard(), False, None),
    ("B34/S3456/C4 Moore(4)", (700, 64), (R, P), [0, 1, 2, 3], M.Stochastic(0.5), False, None),
    ("B45/S345/C5 Moore(5)", (512, 300), (C(1), R), [0] * 40 + [1, 2, 3, 4], M.Synchronous(), False, None),
]


def life_rule(name):
    if name == "ltl":
        return ltl_rule()
    if name == "ltl6":
        return M.Life(range(30, 40), range(25, 45), neighbourhood=M.Moore(6))
    if name == "ltl7":
        return M.Life(range(40, 52), range(35, 60), neighbourhood=M.Moore(7))
    if name == "conway":
        return M.Life("conway")
    if " Moore(" in name:
        rs, r = name.split(" Moore(")
        nb = M.Moore(int(r.rstrip(")")))
        return M.Generations(rs, neighbourhood=nb) if "/C" in rs else M.Life(rs, neighbourhood=nb)
    return M.Generations(name)


@pytest.mark.parametrize("k", range(len(LIFE_CASES)))
def test_life_and_generations_gated(oracle, k):
    name, dims, edges, vals, scheme, masked, tune = LIFE_CASES[k]
    u = rng_state(100 + k, dims, vals)
    prob = CAProblem(life_rule(name), u, (2, 26), space=M.Lattice(dims, edges), seed=11 + k,
                     mask=mask_of(dims, k) if masked else None)
    be = HIPBackend(tune=tune)
    sol = solve(prob, scheme, backend=be, saveat=6)
    assert_same(sol, oracle_saves(oracle, prob, scheme, sol.t))


FOREST_CASES = [
    # dims, edges, p, f, scheme, masked, values
    ((1024, 600), (R, R), 0.03, 1e-3, M.Stochastic(0.5), False, [0, 1, 2]),
    ((1024, 256), (P, P), F32(0.2), 0.01, M.Stochastic(F32(0.6)), True, [0, 1, 2]),
    ((1000, 37), (C(1), P), 0.1, F32(0.05), M.Checkerboard(), True, [0, 1, 2]),
    ((2048, 48), (R, C(2)), F32(0.3), F32(0.02), M.Checkerboard(), False, [0, 1]),
    ((512, 120), (P, R), 0.2, 0.05, M.Synchronous(), True, [0, 1, 2]),
    ((250, 33), (R, R), 0.5, 0.5, M.Stochastic(0.9), True, [0, 1, 2]),
]


@pytest.mark.parametrize("k", range(len(FOREST_CASES)))
def test_forest_gated(oracle, k):
    dims, edges, p, f, scheme, masked, vals = FOREST_CASES[k]
    u = rng_state(200 + k, dims, vals)
    prob = CAProblem(M.ForestFire(), u, (5, 35), {"p": p, "f": f}, space=M.Lattice(dims, edges), seed=21 + k,
                     mask=mask_of(dims, 50 + k) if masked else None)
    sol = solve(prob, scheme, backend=HIPBackend(), saveat=10)
    assert_same(sol, oracle_saves(oracle, prob, scheme, sol.t))


def advance_raw(prob, scheme, u, t0, nsteps):
    """The state after nsteps from t0 and state u, set through the state API (no
    state validation), and the state's kernel names from a dry replay."""
    be = HIPBackend()
    st = be.allocate_state(prob, scheme)
    try:
        be.set_state(st, u)
        be.advance(st, prob, scheme, prob.p, t0, nsteps)
        return be.current_state(st), be.check_launches(prob, scheme)
    finally:
        be.free_state(st)


@pytest.mark.parametrize("scheme", [M.Synchronous(), M.Stochastic(0.5), M.Checkerboard()],
                         ids=["sync", "stochastic", "checkerboard"])
def test_cells_outside_the_states(oracle, scheme):
    """Int8 values no rule writes (negative, past C, 127) are kept or advanced
    exactly as the oracle's cell rule does."""
    odd = [-128, -5, -1, 0, 1, 2, 3, 5, 100, 127]
    for k, (rule, dims, p) in enumerate([(M.Generations("star_wars"), (1024, 64), None),
                                         (M.Generations("B2/S/C130"), (300, 50), None),
                                         (M.Life("B3/S23"), (512, 40), None),
                                         (M.ForestFire(), (1024, 64), {"p": 0.3, "f": 0.1})]):
        u = rng_state(300 + k, dims, odd)
        prob = CAProblem(rule, np.zeros(dims, np.int8), 4, p, space=M.Lattice(dims, (R, P)), seed=5,
                         mask=mask_of(dims, 70 + k))
        got, kernels = advance_raw(prob, scheme, u, 3, 4)
        assert set(kernels) <= {"k_life_gated", "k_forest_gated"}, kernels
        d, keep = desc_for(prob, scheme, field_method=0)
        want = oracle.advance_plain(d, u, 3, 4, params_for(prob))
        bad = np.argwhere(got != want)
        assert not len(bad), f"{rule.__class__.__name__}: {len(bad)} cells differ, first {bad[0].tolist()}"


@pytest.mark.parametrize("rule", ["brians_brain", "star_wars"])
def test_negative_cells_keep_the_exact_rule(oracle, rule):
    """Generations of 3 or 4 states moves a negative cell up by one a step, so a
    sweep over every cell does not leave a lattice of cells below 4 only: the
    walk must keep its exact byte tests on such a state for every step."""
    dims = (1024, 64)
    u = rng_state(400, dims, [-7, -3, -1, 0, 1, 2, 3, 5])
    prob = CAProblem(M.Generations(rule), np.zeros(dims, np.int8), 6, space=M.Lattice(dims, (P, R)))
    got, kernels = advance_raw(prob, M.Synchronous(), u, 0, 6)
    assert set(kernels) == {"k_life_gated"}, kernels
    d, keep = desc_for(prob, M.Synchronous(), field_method=0)
    want = oracle.advance_plain(d, u, 0, 6, params_for(prob))
    assert (want < 0).any()  # negative cells are left after six steps
    bad = np.argwhere(got != want)
    assert not len(bad), f"{len(bad)} cells differ, first {bad[0].tolist()}"
