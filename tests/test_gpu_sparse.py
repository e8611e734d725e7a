"""The sparse sweep (cah_state_sparse: SparseBackend, src/sparse.jl) equals the
dense sweep, and the oracle, at every saved state: partial tiles on periodic
axes, halos wider than a tile, Reflecting and Constant edges (a Constant edge
that is not the background keeps its tiles active), Generations, the forest
fire burning through a full forest, Stochastic gates, masks, 1-D lattices, cell
writes between steps; and it refuses rules that are not quiescent."""
import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend, HIPSparseBackend, make_desc, make_params
from cahip.problem import CAProblem, solve
from helpers import ltl_rule

pytestmark = pytest.mark.gpu


def oracle_saves(oracle, prob, scheme, times):
    d, keep = make_desc(prob, scheme)
    cp = make_params(prob, prob.p)
    out, u, t = [prob.u0], prob.u0, times[0]
    for ts in times[1:]:
        u = oracle.advance_plain(d, u, t, ts - t, cp)
        out.append(u)
        t = ts
    return out


def same(prob, bg, tile, scheme=None, saveat=10, oracle=None):
    scheme = M.Synchronous() if scheme is None else scheme
    sb = HIPSparseBackend(bg, tilesize=tile)
    sp = solve(prob, scheme, backend=sb, saveat=saveat)
    dn = solve(prob, scheme, backend=HIPBackend(), saveat=saveat)
    for k, (a, b) in enumerate(zip(sp.u, dn.u)):
        assert np.array_equal(a, b), f"save {k} (t={sp.t[k]}): sparse and dense differ in {(a != b).sum()} cells"
    if oracle is not None:
        for a, b in zip(sp.u, oracle_saves(oracle, prob, scheme, sp.t)):
            assert np.array_equal(a, b)
    return sp


def blob(dims, centre, size, p, seed):
    rng = np.random.default_rng(seed)
    u = np.zeros(dims, np.int8)
    x, y = centre
    u[x:x + size[0], y:y + size[1]] = (rng.random(size) < p).astype(np.int8)
    return u


def test_life_partial_tiles_on_a_torus(oracle):
    u = np.zeros((1000, 900), np.int8)
    u[500:503, 450:453] = np.array([[0, 1, 1], [1, 1, 0], [0, 1, 0]], np.int8)  # R-pentomino
    prob = CAProblem(M.Life(), u, 300)
    sol = same(prob, 0, 32, saveat=50, oracle=oracle)
    assert (sol.u[-1] == 1).sum() > 50


def test_halo_wider_than_a_tile():
    prob = CAProblem(ltl_rule(), blob((517, 389), (240, 180), (40, 40), 0.5, 1), 60)
    same(prob, 0, 4)
    same(prob, 0, 7)


def test_wrap_across_a_partial_tile():
    # a pattern on the seam of a torus whose sides are not tile multiples
    u = blob((203, 157), (195, 150), (8, 7), 0.6, 2)
    prob = CAProblem(M.Life(range(3, 4), range(2, 4), neighbourhood=M.Moore(3)), u, 40)
    same(prob, 0, 16, saveat=5)


def test_generations_reflecting(oracle):
    u = blob((300, 260), (0, 0), (60, 50), 0.4, 3)
    u[u == 1] = np.random.default_rng(3).choice(np.array([1, 2], np.int8), (u == 1).sum())
    prob = CAProblem(M.Generations("brians_brain"), u, 80, space=M.Grid(300, 260, boundary=M.Reflecting()))
    same(prob, 0, 16, oracle=oracle)


@pytest.mark.parametrize("tile", [8, 32, 100])
def test_constant_edge_not_background(tile):
    u = blob((200, 150), (90, 70), (10, 10), 0.5, 4)
    space = M.Grid(200, 150, boundary=(M.Constant(1), M.Reflecting()))
    prob = CAProblem(M.Life("B3/S23"), u, 50, space=space)
    same(prob, 0, tile)
    space = M.Grid(200, 150, boundary=(M.Periodic(), M.Constant(1)))
    prob = CAProblem(M.Life("B3/S23", neighbourhood=M.Moore(2)), u, 30, space=space)
    same(prob, 0, tile)


def test_forest_fire_burns_through_a_full_forest(oracle):
    u = np.ones((300, 280), np.int8)
    u[150, 140] = 2
    prob = CAProblem(M.ForestFire(), u, 120, {"p": 0.01, "f": 0.0}, seed=5,
                     space=M.Grid(300, 280, boundary=M.Reflecting()))
    sol = same(prob, 1, 16, saveat=20, oracle=oracle)
    assert (sol.u[-1] == 0).sum() > 1000


def test_stochastic_gate_and_mask():
    u = blob((256, 256), (100, 100), (50, 50), 0.5, 6)
    mask = np.ones((256, 256), bool)
    mask[110:130, :] = False
    prob = CAProblem(M.Life(), u, 60, seed=3, mask=mask)
    for scheme in (M.Stochastic(0.5), M.Stochastic(np.float32(0.3))):
        same(prob, 0, 32, scheme=scheme)


def test_one_dimensional():
    u = np.zeros(5000, np.int8)
    u[2500:2520] = np.random.default_rng(7).integers(0, 2, 20).astype(np.int8)
    for b in (M.Periodic(), M.Reflecting(), M.Constant(0)):
        prob = CAProblem(M.Life("B1/S1", neighbourhood=M.Moore(2)), u, 100, space=M.Grid(5000, boundary=b))
        same(prob, 0, 16)


def test_cell_writes_between_steps():
    u = blob((400, 400), (200, 200), (20, 20), 0.5, 8)
    prob = CAProblem(M.Life(), u, 30)
    sb, db = HIPSparseBackend(0, tilesize=32), HIPBackend()
    ss, ds = sb.allocate_state(prob, M.Synchronous()), db.allocate_state(prob, M.Synchronous())
    try:
        for be, st in ((sb, ss), (db, ds)):
            be.advance(st, prob, M.Synchronous(), None, 0, 10)
            for i, j in ((5, 5), (6, 5), (7, 5), (390, 20)):  # a blinker far from the blob
                be.write_cell(st, (i, j), 1)
            be.advance(st, prob, M.Synchronous(), None, 10, 20)
        assert np.array_equal(sb.current_state(ss), db.current_state(ds))
        n = sb.stored_tiles(ss)
        assert 0 < n < (400 // 32 + 1) ** 2
    finally:
        sb.free_state(ss)
        db.free_state(ds)


def test_stored_tiles_follow_the_pattern():
    u = np.zeros((1024, 1024), np.int8)
    u[10, 10] = 1  # a lone cell: it dies, and nothing is left
    prob = CAProblem(M.Life(), u, 3)
    sb = HIPSparseBackend(0, tilesize=32)
    st = sb.allocate_state(prob, M.Synchronous())
    try:
        assert sb.stored_tiles(st) == 1
        sb.advance(st, prob, M.Synchronous(), None, 0, 1)
        assert sb.stored_tiles(st) == 0
        assert not sb.current_state(st).any()
    finally:
        sb.free_state(st)


def test_refuses_rules_that_are_not_quiescent():
    u = np.zeros((64, 64), np.int8)
    with pytest.raises(M.ArgumentError):
        solve(CAProblem(M.Life("B0/S23"), u, 2), backend=HIPSparseBackend(0))
    with pytest.raises(M.ArgumentError):  # an empty forest grows trees
        solve(CAProblem(M.ForestFire(), u, 2, {"p": 0.01, "f": 0.0}), backend=HIPSparseBackend(0))
    with pytest.raises(M.ArgumentError):
        solve(CAProblem(M.FieldEpidemic(M.UniformKernel(1)), u, 2, {"beta": 1.0}), backend=HIPSparseBackend(0))
    with pytest.raises(M.UnsupportedCombinationError):
        solve(CAProblem(M.Life(), u, 2), M.Checkerboard(), backend=HIPSparseBackend(0))
