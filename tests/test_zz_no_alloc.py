"""cah_batch_counts_noalloc, the batch counts that allocate no device memory
(their counters live in a fixed device array): the same counts and
agreements as cah_batch_counts on two calls in a row, the second on a smaller
batch of smaller lattices; the exact-histogram path for values outside the
rule; and the refusal of a batch beyond its maximum, before any launch."""
import ctypes

import numpy as np
import pytest

from cahip import _lib as L
from cahip import model as M
from cahip.backend import HIPBackend, check, make_desc, make_params
from cahip.problem import CAProblem
from helpers import forest_problem, rng_state

pytestmark = pytest.mark.gpu


def make_batch(be, probs, states=None):
    descs, keeps = zip(*[make_desc(p, M.Synchronous()) for p in probs])
    arr = (L.CahDesc * len(descs))(*descs)
    h = ctypes.c_void_p()
    check(be.lib.cah_batch_create(be.ctx(), len(descs), arr, ctypes.byref(h)))
    for i, p in enumerate(probs):
        u = np.asfortranarray(p.u0 if states is None else states[i])
        check(be.lib.cah_batch_set(h, i, u.ctypes.data, u.nbytes))
    return h


def both(be, h, n, obs=None):
    """(counts, equal) from cah_batch_counts and from cah_batch_counts_noalloc."""
    out = []
    for fn in (be.lib.cah_batch_counts, be.lib.cah_batch_counts_noalloc):
        counts = np.full((n, 256), -1, np.int64)
        eq = np.full(n, -1, np.int64)
        check(fn(h, obs.handle if obs is not None else None, counts.ctypes.data,
                 eq.ctypes.data if obs is not None else None))
        out.append((counts, eq))
    return out


def test_noalloc_counts_equal_batch_counts():
    be = HIPBackend()
    # first call: 9 forest fires of ragged sizes after 25 steps, with an observed map
    dims1 = [(300, 211), (150, 97), (1024, 768), (64, 64), (517, 60), (33, 1030), (800, 800), (97, 3), (2000, 100)]
    probs1 = [forest_problem(d, p=0.1, f=0.01, seed=i + 1, u0=rng_state(i, d, [0, 1, 1, 2]), tspan=25)
              for i, d in enumerate(dims1)]
    h1 = make_batch(be, probs1)
    obs = None
    try:
        check(be.lib.cah_batch_advance(h1, 0, 25, ctypes.byref(make_params(probs1[0], probs1[0].p)), 1))
        (c_ref, _), (c_na, _) = both(be, h1, len(probs1))
        assert np.array_equal(c_ref, c_na)
        assert all(c_na[i].sum() == np.prod(d) for i, d in enumerate(dims1))
        # agreement with an observed map needs one lattice shape: a batch of 6 of (300, 211)
        same = [forest_problem((300, 211), p=0.1, f=0.01, seed=i + 1, u0=rng_state(20 + i, (300, 211), [0, 1, 2]),
                               tspan=10) for i in range(6)]
        h2 = make_batch(be, same)
        try:
            check(be.lib.cah_batch_advance(h2, 0, 10, ctypes.byref(make_params(same[0], same[0].p)), 1))
            obs = be.state_like(same[0], M.Synchronous(), rng_state(99, (300, 211), [0, 1, 2]))
            (c_ref, e_ref), (c_na, e_na) = both(be, h2, len(same), obs)
            assert np.array_equal(c_ref, c_na) and np.array_equal(e_ref, e_na)
            assert (e_na > 0).all()
        finally:
            be.lib.cah_batch_destroy(h2)
    finally:
        be.lib.cah_batch_destroy(h1)
        if obs is not None:
            be.free_state(obs)
    # second call in a row: a smaller batch of smaller lattices (the array keeps
    # the last call's contents, which must not leak into this one)
    dims3 = [(40, 30), (17, 5)]
    probs3 = [forest_problem(d, p=0.3, f=0.05, seed=7 + i, u0=rng_state(40 + i, d, [0, 1, 2]), tspan=5)
              for i, d in enumerate(dims3)]
    h3 = make_batch(be, probs3)
    try:
        check(be.lib.cah_batch_advance(h3, 0, 5, ctypes.byref(make_params(probs3[0], probs3[0].p)), 0))
        (c_ref, _), (c_na, _) = both(be, h3, len(probs3))
        assert np.array_equal(c_ref, c_na)
        assert [c_na[i].sum() for i in range(2)] == [1200, 85]
    finally:
        be.lib.cah_batch_destroy(h3)


def test_noalloc_counts_values_outside_the_rule():
    """Lattices holding values the rule never makes (3, -1, 127) take the exact
    256-value histogram; a Generations batch advances member by member."""
    be = HIPBackend()
    dims = [(130, 70), (64, 33)]
    probs = [forest_problem(d, seed=1 + i) for i, d in enumerate(dims)]
    states = [rng_state(60 + i, d, [0, 1, 2, 3, -1, 127]) for i, d in enumerate(dims)]
    h = make_batch(be, probs, states)
    try:
        (c_ref, _), (c_na, _) = both(be, h, 2)
        assert np.array_equal(c_ref, c_na)
        for i, u in enumerate(states):
            vals, cnt = np.unique(u.astype(np.int64) & 0xFF, return_counts=True)
            assert dict(zip(vals.tolist(), cnt.tolist())) == {v: int(c) for v, c in enumerate(c_na[i]) if c}
    finally:
        be.lib.cah_batch_destroy(h)
    gens = [CAProblem(M.Generations("brians_brain"), rng_state(70 + i, (90, 80), [0, 1, 2]), 6, seed=i)
            for i in range(3)]
    h = make_batch(be, gens)
    try:
        check(be.lib.cah_batch_advance(h, 0, 6, ctypes.byref(make_params(gens[0], None)), 0))
        (c_ref, _), (c_na, _) = both(be, h, 3)
        assert np.array_equal(c_ref, c_na)
    finally:
        be.lib.cah_batch_destroy(h)


def test_noalloc_counts_refuses_too_many_trajectories():
    be = HIPBackend()
    n = L.COUNTS_NOALLOC_MAX + 1
    probs = [CAProblem(M.Life(), np.zeros((4, 4), np.int8), 1, seed=i) for i in range(n)]
    h = make_batch(be, probs)
    try:
        counts = np.zeros((n, 256), np.int64)
        rc = be.lib.cah_batch_counts_noalloc(h, None, counts.ctypes.data, None)
        assert rc == L.CAH_ERR_ARGUMENT and "at most 512" in L.last_error()
        assert not counts.any()
    finally:
        be.lib.cah_batch_destroy(h)
