"""Cell counts and map agreement counted on the device (cah_state_counts,
cah_batch_counts) against numpy on the same states: state_counts
(src/measures.jl:14-20) and cohens_kappa (src/evaluation.jl:30-50) exactly,
for ragged row lengths, 1-D lattices, values the rule never produces (the full
histogram path) and batched ensembles."""
import numpy as np
import pytest

from cahip import measures as MS
from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, final_state, solve
from helpers import forest_problem

pytestmark = pytest.mark.gpu


def host_counts(u):
    return MS.state_counts(u)


@pytest.mark.parametrize("dims", [(1, 1), (15, 3), (16, 1), (17, 9), (1023, 5), (1025, 7), (2049, 33),
                                  (300, 257), (4096, 64)])
@pytest.mark.parametrize("rule,values", [(M.Life(), 2), (M.ForestFire(), 3), (M.Generations("B2/S/C7"), 7)])
def test_counts_match_numpy(backend, dims, rule, values):
    rng = np.random.default_rng(dims[0] * 131 + dims[1])
    u = rng.integers(0, values, dims).astype(np.int8)
    o = rng.integers(0, values, dims).astype(np.int8)
    o[: dims[0] // 2] = u[: dims[0] // 2]  # agree on part of the map
    prob = CAProblem(rule, u, 1, {"p": 0.1, "f": 0.01} if isinstance(rule, M.ForestFire) else None)
    st = backend.allocate_state(prob, M.Synchronous())
    ob = backend.state_like(prob, M.Synchronous(), o)
    try:
        assert backend.state_counts(st) == host_counts(u)
        assert backend.cohens_kappa(st, ob) == MS.cohens_kappa(u, o)
        counts, eq = backend._counts(st, ob)
        assert eq == int(np.count_nonzero(u == o))
        assert counts.sum() == u.size
    finally:
        backend.free_state(st)
        backend.free_state(ob)


def test_one_dimensional_lattice(backend):
    u = np.random.default_rng(3).integers(0, 2, 5000).astype(np.int8)
    prob = CAProblem(M.Life(neighbourhood=M.Moore(1)), u, 1, space=M.Torus(5000))
    st = backend.allocate_state(prob, M.Synchronous())
    try:
        assert backend.state_counts(st) == host_counts(u)
    finally:
        backend.free_state(st)


def test_values_outside_the_rule(backend):
    # Life never makes -7, 5 or 127, but a state set by the caller may hold them:
    # the full 256-value histogram counts every cell
    rng = np.random.default_rng(8)
    u = rng.integers(0, 2, (333, 71)).astype(np.int8)
    u[rng.random(u.shape) < 0.01] = -7
    u[5, 5], u[6, 6] = 5, 127
    prob = CAProblem(M.Life(), np.zeros_like(u), 1)
    st = backend.allocate_state(prob, M.Synchronous())
    try:
        backend.set_state(st, u)
        got = backend.state_counts(st)
        assert got == host_counts(u)
        assert got[-7] > 0 and got[5] == 1 and got[127] == 1
    finally:
        backend.free_state(st)


def test_counts_after_steps(backend):
    # counts of an advanced forest fire equal those of its downloaded state
    prob = forest_problem((1200, 700), p=0.05, f=0.001, tspan=40)
    st = backend.allocate_state(prob, M.Synchronous())
    try:
        backend.advance(st, prob, M.Synchronous(), prob.p, 0, 40)
        u = backend.current_state(st)
        assert backend.state_counts(st) == host_counts(u)
        assert set(backend.state_counts(st)) <= {0, 1, 2}
    finally:
        backend.free_state(st)


def test_kappa_dims_must_match(backend):
    a = CAProblem(M.Life(), np.zeros((40, 30), np.int8), 1)
    b = CAProblem(M.Life(), np.zeros((30, 40), np.int8), 1)
    sa, sb = backend.allocate_state(a, M.Synchronous()), backend.allocate_state(b, M.Synchronous())
    try:
        with pytest.raises(M.StateShapeError):
            backend.cohens_kappa(sa, sb)
    finally:
        backend.free_state(sa)
        backend.free_state(sb)


def test_batch_counts_and_kappa():
    """An ensemble of forest fires reduced on the device, as an output_func that
    keeps only state counts or 1 - kappa against an observed map would
    (src/calibration.jl:76)."""
    be = HIPBackend()
    dims = (517, 300)
    probs = [forest_problem(dims, p=0.04, f=0.002, seed=s, tspan=25) for s in range(1, 13)]
    finals = be.solve_batch(probs, M.Synchronous())
    counts = be.solve_batch(probs, M.Synchronous(), output="counts")
    assert counts == [host_counts(u) for u in finals]
    observed = finals[0]
    kappas = be.solve_batch(probs, M.Synchronous(), output="kappa", observed=observed)
    assert kappas == [MS.cohens_kappa(u, observed) for u in finals]
    assert kappas[0] == 1.0
    # ragged ensemble: counts need no common dims
    rag = [forest_problem((100 + 37 * s, 64 + 11 * s), seed=s, tspan=10) for s in range(1, 6)]
    fin = be.solve_batch(rag, M.Synchronous())
    assert be.solve_batch(rag, M.Synchronous(), output="counts") == [host_counts(u) for u in fin]
    with pytest.raises(M.StateShapeError):
        be.solve_batch(rag, M.Synchronous(), output="kappa", observed=fin[0])


def test_batch_saveat_series():
    """solve_batch with saveat: per trajectory, the states / counts / kappa at
    every saved time equal those of the trajectory's own solve at those times
    (counts and kappa counted on the device between the batch's advances)."""
    be = HIPBackend()
    dims = (300, 211)
    probs = [forest_problem(dims, p=0.05, f=0.003, seed=s, tspan=(2, 40)) for s in range(1, 7)]
    saveat = [5, 17, 18, 33]
    times = [2, 5, 17, 18, 33, 40]
    alone = [solve(p, M.Synchronous(), backend=be, saveat=saveat) for p in probs]
    assert all(sol.t == times for sol in alone)
    states = be.solve_batch(probs, M.Synchronous(), saveat=saveat)
    counts = be.solve_batch(probs, M.Synchronous(), output="counts", saveat=saveat)
    observed = alone[0].u[-1]
    kappas = be.solve_batch(probs, M.Synchronous(), output="kappa", observed=observed, saveat=saveat)
    for i, sol in enumerate(alone):
        assert len(states[i]) == len(counts[i]) == len(kappas[i]) == len(times)
        for k in range(len(times)):
            assert np.array_equal(states[i][k], sol.u[k]), (i, times[k])
            assert counts[i][k] == host_counts(sol.u[k]), (i, times[k])
            assert kappas[i][k] == MS.cohens_kappa(sol.u[k], observed), (i, times[k])
    # a step size: every 10 steps from t0, and t1
    every = be.solve_batch(probs[:2], M.Synchronous(), output="counts", saveat=10)
    sol = solve(probs[0], M.Synchronous(), backend=be, saveat=10)
    assert sol.t == [2, 12, 22, 32, 40]
    assert every[0] == [host_counts(u) for u in sol.u]
    # the ensemble form, batched and one integrator per trajectory
    from cahip.problem import EnsembleProblem, solve_ensemble
    ep = EnsembleProblem(probs[0])
    b = solve_ensemble(ep, backend=be, trajectories=4, saveat=saveat)
    u = solve_ensemble(ep, backend=be, trajectories=4, batched=False, saveat=saveat)
    assert all(len(x) == len(times) and all(np.array_equal(p, q) for p, q in zip(x, y)) for x, y in zip(b, u))


def test_full_size_forest_counts(backend):
    """16384^2 forest fire after 30 steps: the counts add up to every cell and
    equal numpy's on the downloaded state."""
    prob = forest_problem((16384, 16384), tspan=30)
    st = backend.allocate_state(prob, M.Synchronous())
    try:
        backend.advance(st, prob, M.Synchronous(), prob.p, 0, 30)
        got = backend.state_counts(st)
        assert sum(got.values()) == 16384 * 16384
        u = backend.current_state(st)
        assert got == host_counts(u)
    finally:
        backend.free_state(st)
