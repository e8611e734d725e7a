"""The counting kernels that the band walks leave, against the CPU oracle bit for
bit at every save (src/backends.jl:444-455, src/schemes.jl:74-93,
src/rules.jl:377-389, :446-456):

- k_box_i8, Life and Generations with Moore(R <= 32) or VonNeumann(r <= 8)
  where no band walk runs (Moore radii above 6, more than two count
  intervals, every von Neumann radius): counts as window sums of prefix
  counts;
- k_generic_i8, everything else on Int8 lattices (1-D lattices, von Neumann
  radii above 8, explicit offsets): four cells a thread, counted a dword at a
  time, its byte counters flushed every 255 offsets.

The lattices are wide enough for interior tiles (filled a dword at a time)
beside tiles at every edge kind (filled through the boundary map), or smaller
than one tile (read through several periods); first sides that are not a
multiple of 4 (the row's last cells written one by one); masks and the gated
schemes.

test_cases_take_their_kernel runs without a device: the host-side launch
checks replay every launch of each case and name its kernel.
"""
import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, solve
from helpers import rng_state
from test_gpu_parity import assert_same, oracle_saves

P, R, C = M.Periodic(), M.Reflecting(), M.Constant

SPARSE = [0] * 7 + [1]   # one cell in eight alive: Moore(8) counts stay in the rule's 0..63
SPARSER = [0] * 99 + [1]  # one in a hundred: Moore(32) counts (of 4224 neighbours) too

BOX, GEN = "k_box_i8", "k_generic_i8"

CASES = [
    # name, rule, space, values, scheme, masked, kernel
    ("vn2_torus", M.Life("B2/S13", neighbourhood=M.VonNeumann(2)), M.Torus(300, 200), [0, 1],
     M.Synchronous(), False, BOX),
    ("vn2_mixed_edges", M.Life("B3/S2345", neighbourhood=M.VonNeumann(2)),
     M.Lattice((517, 61), (R, C(1))), [0, 1], M.Stochastic(0.5), True, BOX),
    ("vn3_f32_gate", M.Life("B3/S2345", neighbourhood=M.VonNeumann(3)),
     M.Lattice((259, 140), (C(0), P)), [0, 1], M.Stochastic(np.float32(0.6)), False, BOX),
    ("vn1_majority_checkerboard", M.Life("B5678/S45678", neighbourhood=M.VonNeumann(1)),
     M.Lattice((230, 130), (R, P)), [0, 1], M.Checkerboard(), False, BOX),
    ("vn8_reflecting", M.Life(range(20, 40), range(15, 45), neighbourhood=M.VonNeumann(8)),
     M.Lattice((333, 150), (R, R)), [0, 1], M.Synchronous(), True, BOX),
    ("moore9_checkerboard", M.Life(range(25, 40), range(20, 45), neighbourhood=M.Moore(9)),
     M.Lattice((250, 130), (P, R)), SPARSE, M.Checkerboard(), False, BOX),
    ("moore6_four_intervals", M.Life([30, 32, 34, 36], [25, 28, 31, 40], neighbourhood=M.Moore(6)),
     M.Lattice((270, 110), (R, C(1))), [0, 0, 0, 0, 1], M.Synchronous(), False, BOX),
    ("moore8_masked", M.Life(range(20, 30), range(15, 35), neighbourhood=M.Moore(8)), M.Torus(200, 150), SPARSE,
     M.Synchronous(), True, BOX),
    ("moore8_small_torus", M.Life(range(20, 30), range(15, 35), neighbourhood=M.Moore(8)), M.Torus(45, 37), SPARSE,
     M.Synchronous(), False, BOX),
    ("moore32_constant", M.Life(range(30, 50), range(25, 60), neighbourhood=M.Moore(32)),
     M.Lattice((130, 70), (C(0), P)), SPARSER, M.Stochastic(0.8), False, BOX),
    # every cell's window covers nearly the whole torus (some cells twice): the
    # first step's births mark where the count reaches 36
    ("moore32_narrow_torus", M.Life(range(36, 41), range(36, 41), neighbourhood=M.Moore(32)),
     M.Torus(90, 50), SPARSER, M.Synchronous(), False, BOX),
    ("life_three_intervals", M.Life([1, 3, 5, 6], [2, 4, 7, 8], neighbourhood=M.Moore(2)),
     M.Lattice((333, 97), (R, R)), [0, 1], M.Synchronous(), False, BOX),
    ("replicator_moore1", M.Life("B1357/S1357"), M.Lattice((190, 70), (C(1), P)), [0, 1], M.Synchronous(),
     False, BOX),
    ("generations_vn2", M.Generations("B2/S13/C4", neighbourhood=M.VonNeumann(2)),
     M.Lattice((402, 75), (C(2), R)), [0, 1, 2, 3], M.Checkerboard(), True, BOX),
    ("generations_moore9", M.Generations("B45/S345/C5", neighbourhood=M.Moore(9)),
     M.Lattice((300, 140), (P, C(3))), [0] * 40 + [1, 2, 3, 4], M.Stochastic(0.7), True, BOX),
    ("one_dimensional", M.Life("B1/S1", neighbourhood=M.Moore(2)), M.Grid(5001, boundary=C(1)), [0, 1],
     M.Stochastic(0.7), False, GEN),
    ("vn12_flush", M.Life(range(30, 45), range(25, 50), neighbourhood=M.VonNeumann(12)), M.Torus(200, 150), SPARSE,
     M.Synchronous(), True, GEN),
    ("offsets", M.Life("B1/S12", neighbourhood=M.Offsets((2, 1), (-3, 0), (0, 4), (1, -1))),
     M.Lattice((260, 90), (R, C(1))), [0, 1], M.Synchronous(), False, GEN),
]


def problem(k):
    name, rule, space, vals, scheme, masked, _ = CASES[k]
    u = rng_state(400 + k, space.dims, vals)
    mask = None
    if masked:
        mask = np.random.default_rng(500 + k).random(space.dims) >= 0.15
    return CAProblem(rule, u, (3, 23), space=space, seed=31 + k, mask=mask), scheme


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_cases_take_their_kernel(k):
    prob, scheme = problem(k)
    rep = HIPBackend().check_launches(prob, scheme)
    assert set(rep) == {CASES[k][6]}, rep


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_counting_kernels_against_the_oracle(oracle, k):
    prob, scheme = problem(k)
    sol = solve(prob, scheme, backend=HIPBackend(), saveat=1)
    assert_same(sol, oracle_saves(oracle, prob, scheme, sol.t, halo=False))
