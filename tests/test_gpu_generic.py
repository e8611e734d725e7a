"""The general Int8 sweep (k_generic_i8) against the CPU oracle, bit for bit at
every save, on lattices wide enough for its interior tiles (filled a dword at
a time) beside tiles at every edge kind (filled through the boundary map):
von Neumann radii 2 and 3, Moore radii 6 and 8 (168 and 288 neighbours, past
the 255 offsets after which the kernel flushes its byte counters), Life with
three count intervals, Generations with a von Neumann neighbourhood, a 1-D
lattice, first sides that are not a multiple of 4 (the row's last cells
written one by one), masks and the gated schemes (src/backends.jl:444-455,
src/schemes.jl:74-93, src/rules.jl:377-389, :446-456).

test_cases_take_the_general_kernel runs without a device: the host-side
launch checks replay every launch of each case and name its kernel.
"""
import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, solve
from helpers import rng_state
from test_gpu_parity import assert_same, oracle_saves

P, R, C = M.Periodic(), M.Reflecting(), M.Constant

SPARSE = [0] * 7 + [1]  # one cell in eight alive: counts of Moore(8) stay in the rule's 0..63

CASES = [
    # name, rule, space, values, scheme, masked
    ("vn2_torus", M.Life("B2/S13", neighbourhood=M.VonNeumann(2)), M.Torus(300, 200), [0, 1],
     M.Synchronous(), False),
    ("vn2_mixed_edges", M.Life("B3/S2345", neighbourhood=M.VonNeumann(2)),
     M.Lattice((517, 61), (R, C(1))), [0, 1], M.Stochastic(0.5), True),
    ("vn3_f32_gate", M.Life("B3/S2345", neighbourhood=M.VonNeumann(3)),
     M.Lattice((259, 140), (C(0), P)), [0, 1], M.Stochastic(np.float32(0.6)), False),
    ("moore6_checkerboard", M.Life(range(25, 40), range(20, 45), neighbourhood=M.Moore(6)),
     M.Lattice((250, 130), (P, R)), [0, 1], M.Checkerboard(), False),
    ("moore8_flush", M.Life(range(20, 30), range(15, 35), neighbourhood=M.Moore(8)), M.Torus(200, 150), SPARSE,
     M.Synchronous(), True),
    ("life_three_intervals", M.Life([1, 3, 5, 6], [2, 4, 7, 8], neighbourhood=M.Moore(2)),
     M.Lattice((333, 97), (R, R)), [0, 1], M.Synchronous(), False),
    ("generations_vn2", M.Generations("B2/S13/C4", neighbourhood=M.VonNeumann(2)),
     M.Lattice((402, 75), (C(2), R)), [0, 1, 2, 3], M.Checkerboard(), True),
    ("one_dimensional", M.Life("B1/S1", neighbourhood=M.Moore(2)), M.Grid(5001, boundary=C(1)), [0, 1],
     M.Stochastic(0.7), False),
]


def problem(k):
    name, rule, space, vals, scheme, masked = CASES[k]
    u = rng_state(400 + k, space.dims, vals)
    mask = None
    if masked:
        mask = np.random.default_rng(500 + k).random(space.dims) >= 0.15
    return CAProblem(rule, u, (3, 23), space=space, seed=31 + k, mask=mask), scheme


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_cases_take_the_general_kernel(k):
    prob, scheme = problem(k)
    rep = HIPBackend().check_launches(prob, scheme)
    assert set(rep) == {"k_generic_i8"}, rep


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_general_kernel_against_the_oracle(oracle, k):
    prob, scheme = problem(k)
    sol = solve(prob, scheme, backend=HIPBackend(), saveat=5)
    assert_same(sol, oracle_saves(oracle, prob, scheme, sol.t, halo=False))
