"""Shared helpers for the tests: problem/descriptor construction and naive
definitions written from the reference's own test definitions."""
import json
import os

import numpy as np

from cahip import model as M
from cahip.backend import make_desc, make_params
from cahip.problem import CAProblem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def rng_state(seed, dims, values, dtype=np.int8):
    """A seeded random state (numpy's generator; the reference draws its test
    states from Xoshiro, which is not available here)."""
    r = np.random.default_rng(seed)
    return np.asfortranarray(r.choice(np.asarray(values), size=dims).astype(dtype))


def desc_for(prob, scheme=None, field_method=1):
    scheme = M.Synchronous() if scheme is None else scheme
    return make_desc(prob, scheme, field_method)


def params_for(prob, p=None):
    return make_params(prob, prob.p if p is None else p)


def oracle_final(oracle, prob, scheme=None, nthreads=None):
    """The oracle's state at the end of prob's tspan (the threaded ArrayBackend
    restatement, for Int8 rules)."""
    import os
    d, keep = desc_for(prob, scheme, field_method=0)
    t0, t1 = prob.tspan
    nt = nthreads or min(16, os.cpu_count() or 1)
    return oracle.advance_halo(d, prob.u0, t0, t1 - t0, params_for(prob), nt)


def naive_generations(birth, survive, C, u):
    """naive_generations (test/test_rules.jl:23-32): Moore(1) on a torus, C = 2 is Life."""
    m, n = u.shape
    out = np.zeros_like(u)
    for i in range(m):
        for j in range(n):
            live = sum(u[(i + a) % m, (j + b) % n] == 1
                       for a in (-1, 0, 1) for b in (-1, 0, 1) if (a, b) != (0, 0))
            s = u[i, j]
            if s == 0:
                out[i, j] = 1 if live in birth else 0
            elif s == 1:
                out[i, j] = 1 if live in survive else (2 if C > 2 else 0)
            else:
                out[i, j] = s + 1 if s + 1 < C else 0
    return out


def brute_life_torus(u, birth, survive, r):
    """Larger-than-Life by brute force on a torus: the live count over all
    (2r+1)^2 - 1 shifted copies, then the rule's birth/survive sets."""
    cnt = np.zeros(u.shape, dtype=np.int32)
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            if (a, b) != (0, 0):
                cnt += np.roll(np.roll(u == 1, -a, axis=0), -b, axis=1)
    alive = u == 1
    born = np.isin(cnt, list(birth))
    keep = np.isin(cnt, list(survive))
    return np.where(alive, keep, born).astype(np.int8)


def forest_problem(dims, p=0.03, f=1e-4, seed=7, boundary=None, u0=None, tspan=10):
    space = M.Grid(*dims, boundary=M.Reflecting() if boundary is None else boundary)
    u0 = np.zeros(dims, dtype=np.int8) if u0 is None else u0
    return CAProblem(M.ForestFire(), u0, tspan, {"p": p, "f": f}, space=space, seed=seed)


def ltl_rule():
    return M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5))
