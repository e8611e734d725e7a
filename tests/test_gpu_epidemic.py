"""The field epidemic on the GPU against the oracle, state by state, bit for
bit: the kernel-field cases of the reference's conformance battery
(test/backend_conformance.jl:63-104: a Float64 LocalRule with an anisotropic
exponential kernel, a cycling covariate, a mask and mixed edges; a Float32
LocalRule; a Float32 ProbabilisticRule), then the forms, edges, schemes and
sizes around them, batches and the server."""
import math

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend, make_desc, make_params
from cahip.problem import CAProblem, final_state, solve
from cahip.server import Server
from test_epidemic_oracle import windfield_problem

pytestmark = pytest.mark.gpu


def oracle_run(oracle, prob, scheme, times):
    d, keep = make_desc(prob, scheme)
    cp = make_params(prob, prob.p)
    out, u, t = [prob.u0], prob.u0, times[0]
    for ts in times[1:]:
        u = oracle.advance_plain(d, u, t, ts - t, cp)
        out.append(u)
        t = ts
    return out


def assert_run(oracle, prob, scheme=None, saveat=1):
    scheme = M.Synchronous() if scheme is None else scheme
    sol = solve(prob, scheme, backend=HIPBackend(), saveat=saveat)
    ref = oracle_run(oracle, prob, scheme, sol.t)
    for k, (a, b) in enumerate(zip(sol.u, ref)):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"t={sol.t[k]}: {len(bad)} cells differ, first {bad[0].tolist()}: "
                                 f"gpu {a[tuple(bad[0])]} oracle {b[tuple(bad[0])]}")
    return sol


def si0():
    u = np.zeros((30, 26), np.int8)
    u[14, 12] = 1
    return u


def test_conformance_kernel_field_covariate_mask(oracle):
    sol = assert_run(oracle, windfield_problem())
    assert not np.array_equal(sol.u[-1], sol.u[0])


def test_conformance_float32_kernel_field(oracle):
    rule = M.FieldEpidemic(M.AnisotropicKernel(M.GaussianKernel(1.5), 1.0, 1.0), precision=32)
    prob = CAProblem(rule, si0(), 12, {"beta": np.float32(2.0)}, seed=13)
    sol = assert_run(oracle, prob)
    assert (sol.u[-1] == 1).sum() > 1


def test_conformance_stated_float32(oracle):
    rule = M.FieldEpidemic(M.AnisotropicKernel(M.GaussianKernel(1.5), 1.0, 1.0), precision=32, form="stated")
    prob = CAProblem(rule, si0(), 12, {"beta": np.float32(2.0)}, seed=15)
    sol = assert_run(oracle, prob)
    assert (sol.u[-1] == 1).sum() > 1


KERNELS = [M.GaussianKernel(2.0), M.ExponentialKernel(1.2), M.PowerLawKernel(1.5, radius=6), M.UniformKernel(3),
           M.LeniaKernel(7), M.AnisotropicKernel(M.ExponentialKernel(2.0), -0.7, 1.5),
           M.EmpiricalKernel(np.arange(35, dtype=float).reshape(5, 7) % 4)]
EDGES = [M.Periodic(), M.Reflecting(), M.Constant(0), M.Constant(1)]


@pytest.mark.parametrize("k", range(len(KERNELS)))
def test_kernels_and_edges(oracle, k):
    rng = np.random.default_rng(k)
    dims = (157, 131)
    u0 = (rng.random(dims) < 0.02).astype(np.int8)
    e1, e2 = EDGES[k % 4], EDGES[(k + 1) % 4]
    for precision, beta in ((64, 0.8), (32, np.float32(0.8))):
        for form in ("bernoulli", "stated"):
            rule = M.FieldEpidemic(KERNELS[k], precision=precision, form=form, include_centre=k % 2 == 0)
            prob = CAProblem(rule, u0, 6, {"beta": beta}, seed=20 + k, space=M.Grid(*dims, boundary=(e1, e2)))
            assert_run(oracle, prob, saveat=2)


@pytest.mark.parametrize("form", ["bernoulli", "stated"])
@pytest.mark.parametrize("scheme", [M.Synchronous(), M.Stochastic(0.5), M.Stochastic(np.float32(0.7)),
                                    M.Checkerboard()])
def test_sir_schemes(oracle, form, scheme):
    rng = np.random.default_rng(5)
    u0 = rng.choice(np.array([0, 1, 2], np.int8), (96, 200), p=[0.9, 0.08, 0.02])
    rule = M.FieldEpidemic(M.GaussianKernel(1.3), recover=True, form=form)
    for gamma in (0.2, np.float32(0.3)):
        prob = CAProblem(rule, u0, 8, {"beta": 1.1, "gamma": gamma}, seed=31)
        sol = assert_run(oracle, prob, scheme, saveat=4)
        assert (sol.u[-1] == 2).sum() > (u0 == 2).sum()


def test_one_dimensional(oracle):
    u0 = np.zeros(3000, np.int8)
    u0[::97] = 1
    rule = M.FieldEpidemic(M.ExponentialKernel(2.0))
    for b in (M.Periodic(), M.Reflecting(), M.Constant(1)):
        prob = CAProblem(rule, u0, 10, {"beta": 0.9}, seed=2, space=M.Grid(3000, boundary=b))
        assert_run(oracle, prob, saveat=5)


def test_large_lattice(oracle):
    rng = np.random.default_rng(11)
    dims = (1024, 768)
    u0 = (rng.random(dims) < 0.001).astype(np.int8)
    cov = rng.random(dims + (2,))
    rule = M.FieldEpidemic(M.GaussianKernel(2.5), covariate="c")
    prob = CAProblem(rule, u0, 3, {"beta": 3.0}, seed=44, covariates={"c": cov})
    sol = assert_run(oracle, prob, saveat=3)
    assert (sol.u[-1] == 1).sum() > (u0 == 1).sum()


def test_batch_and_server_equal_single_runs():
    rng = np.random.default_rng(17)
    rule = M.FieldEpidemic(M.AnisotropicKernel(M.GaussianKernel(1.5), 0.3, 1.0), recover=True)
    probs = []
    for s in range(6):
        u0 = (rng.random((60 + 9 * s, 50 + 7 * s)) < 0.03).astype(np.int8)
        cov = rng.random(u0.shape + (3,))
        r = M.FieldEpidemic(rule.kernel, recover=True, covariate="c")
        probs.append(CAProblem(r, u0, 9, {"beta": 2.0, "gamma": 0.1}, seed=s, covariates={"c": cov}))
    be = HIPBackend()
    singles = [final_state(solve(p, backend=be, save_everystep=False)) for p in probs]
    outs = be.solve_batch(probs, M.Synchronous())
    assert all(np.array_equal(a, b) for a, b in zip(outs, singles))
    # a budget of about three requests, so they queue
    srv = Server(budget_bytes=1_000_000, chunk=4)
    ids = [srv.submit(p) for p in probs]
    got = {}
    for _ in range(100000):
        srv.tick()
        for rid in srv.poll():
            got[rid] = srv.fetch(rid)
        if len(got) == len(ids):
            break
    for rid, want in zip(ids, singles):
        assert np.array_equal(got[rid], want)
    srv.close()
