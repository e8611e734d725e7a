"""The field epidemic on the CPU: the host mirror's dispersal kernels against
the reference's own checks (test/test_environment.jl:3-43), the oracle's kernel
field against its field values (:46-81), and the oracle's rule against what the
mathematics fixes (no infection at beta = 0, certain infection in reach at a
huge beta, permanent infection without recovery, masks, the covariate's time
slices, the infection rate 1 - exp(-beta * field), the stated form's inversion)."""
import math

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import make_desc, make_params
from cahip.problem import CAProblem

SYNC = M.Synchronous()


def field_of(oracle, u, kernel, space=None, include_centre=True, precision=64):
    u = np.asarray(u, np.int8)
    prob = CAProblem(M.FieldEpidemic(kernel, include_centre=include_centre, precision=precision,
                                     recover=True), u, 1, {"beta": 1.0, "gamma": 0.0}, space=space)
    d, keep = make_desc(prob, SYNC)
    return oracle.epi_field(d, u)


def step(oracle, prob, t0=0, n=1, scheme=SYNC, u=None):
    d, keep = make_desc(prob, scheme)
    return oracle.advance_plain(d, prob.u0 if u is None else u, t0, n, make_params(prob, prob.p))


# ------------------------------------------------------------------ kernels

@pytest.mark.parametrize("k", [M.GaussianKernel(1.5), M.ExponentialKernel(1.0), M.PowerLawKernel(2.0, radius=4),
                               M.UniformKernel(2), M.LeniaKernel(5), M.LeniaKernel(6, peaks=(0.5, 1.0))])
def test_kernel_weights_normalised_within_radius(k):
    o, w = M.kernel_weights(k, M.Torus(20, 20))
    assert abs(w.sum() - 1) < 1e-12 and (w > 0).all() and len(o) == len(w)
    assert all(math.hypot(*x) <= k.radius + 1e-9 for x in o)


def test_kernel_weights_reference_checks():
    o, w = M.kernel_weights(M.UniformKernel(1), M.Torus(5, 5))
    assert set(o) == {(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)}
    o, _ = M.kernel_weights(M.UniformKernel(1), M.Torus(5, 5), include_centre=False)
    assert (0, 0) not in o
    o, w = M.kernel_weights(M.GaussianKernel(1.0, radius=3), M.Torus(9, 9))
    assert w[o.index((0, 0))] > w[o.index((0, 2))]
    o, _ = M.kernel_weights(M.LeniaKernel(4), M.Torus(9, 9))
    assert (0, 0) not in o
    # wind towards the east (+column) favours positive column offsets
    o, w = M.kernel_weights(M.AnisotropicKernel(M.GaussianKernel(1.0, radius=2), 0.0, 2.0), M.Torus(9, 9))
    assert w[o.index((0, 1))] > w[o.index((-1, 0))] > w[o.index((0, -1))]
    o, w = M.kernel_weights(M.AnisotropicKernel(M.GaussianKernel(1.0, radius=2), math.pi / 2, 2.0), M.Torus(9, 9))
    assert w[o.index((-1, 0))] > w[o.index((1, 0))]
    o, w = M.kernel_weights(M.EmpiricalKernel([[0, 1, 0], [2, 4, 2], [0, 1, 0]]), M.Torus(5, 5))
    assert w[o.index((0, 0))] == pytest.approx(0.4)
    with pytest.raises(M.ArgumentError):
        M.GaussianKernel(-1.0)
    with pytest.raises(M.ArgumentError):
        M.EmpiricalKernel([[1, 1], [1, 1]])
    with pytest.raises(M.IncompatibleSpaceError):
        M.kernel_weights(M.EmpiricalKernel(np.ones((3, 3))), M.Torus(9))


def test_kernel_offsets_are_column_major():
    o, _ = M.kernel_weights(M.UniformKernel(1), M.Torus(5, 5))
    assert o == [(0, -1), (-1, 0), (0, 0), (1, 0), (0, 1)]
    o, _ = M.kernel_weights(M.UniformKernel(2), M.Torus(9))
    assert o == [(-2,), (-1,), (0,), (1,), (2,)]


# ------------------------------------------------------------------ the oracle's field

def test_uniform_field_values(oracle):
    u = np.zeros((7, 7), np.int8)
    u[3, 3] = 1
    v = field_of(oracle, u, M.UniformKernel(1))
    assert v[3, 3] == pytest.approx(0.2) and v[2, 3] == pytest.approx(0.2) and v[3, 4] == pytest.approx(0.2)
    assert v[2, 2] == 0
    # beyond a Constant(1) edge cells count as infected
    e = field_of(oracle, np.zeros((5, 5)), M.UniformKernel(1), space=M.Grid(5, 5, boundary=M.Constant(1)))
    assert e[0, 0] == pytest.approx(0.4) and e[0, 2] == pytest.approx(0.2) and e[2, 2] == 0


def test_field_direction_and_empirical(oracle):
    src = np.zeros((9, 9), np.int8)
    src[4, 4] = 1
    w = field_of(oracle, src, M.AnisotropicKernel(M.GaussianKernel(1.0, radius=3), 0.0, 3.0))
    assert w[4, 5] > w[4, 3] and w[3, 4] == pytest.approx(w[5, 4])
    # all mass at offset (-1, -1): propagules travel (-1, -1) from their source, so
    # the whole field sits one cell up and left of the infected cell
    w = field_of(oracle, src, M.EmpiricalKernel([[1, 0, 0], [0, 0, 0], [0, 0, 0]]))
    assert w[3, 3] == pytest.approx(1) and w.sum() == pytest.approx(1)


def test_field32_is_the_float32_sum(oracle):
    rng = np.random.default_rng(1)
    u = (rng.random((30, 26)) < 0.3).astype(np.int8)
    k = M.AnisotropicKernel(M.GaussianKernel(1.5), 1.0, 1.0)
    f64 = field_of(oracle, u, k)
    f32 = field_of(oracle, u, k, precision=32)
    assert np.abs(f64 - f32).max() < 1e-6 and not np.array_equal(f64, f32)


# ------------------------------------------------------------------ the oracle's rule

def windfield_problem(beta=4.0, seed=9, tspan=12):
    """The conformance case "kernel field, cycling covariate and mask"
    (test/backend_conformance.jl:63-71, :162-165), on the build's inputs."""
    rng = np.random.default_rng(2024)
    u0 = np.zeros((30, 26), np.int8)
    u0[14, 12] = 1
    suit = rng.random((30, 26, 3))
    mask = np.ones((30, 26), bool)
    mask[4:10, 4:20] = False
    rule = M.FieldEpidemic(M.AnisotropicKernel(M.ExponentialKernel(1.5), math.pi / 4, 1.0),
                           covariate="suitability")
    return CAProblem(rule, u0, tspan, {"beta": beta}, seed=seed, mask=mask,
                     covariates={"suitability": suit},
                     space=M.Grid(30, 26, boundary=(M.Reflecting(), M.Constant(0))))


def test_beta_zero_and_huge(oracle):
    prob = windfield_problem()
    assert np.array_equal(step(oracle, prob.remake(p={"beta": 0.0}), n=5), prob.u0)
    # a huge beta infects every unmasked susceptible cell with an infected cell in reach
    big = prob.remake(p={"beta": 1e12})
    d, keep = make_desc(big, SYNC)
    f = oracle.epi_field(d, big.u0)
    cov0 = big.covariates["suitability"][:, :, 0]
    want = np.where((f > 0) & (cov0 > 0) & big.mask, 1, big.u0).astype(np.int8)
    assert np.array_equal(step(oracle, big), want)


def test_infection_is_permanent_and_masks_hold(oracle):
    prob = windfield_problem(tspan=12)
    u = prob.u0
    for t in range(12):
        v = step(oracle, prob, t0=t, u=u)
        assert ((u == 1) <= (v == 1)).all()          # no recovery
        assert np.array_equal(v[~prob.mask], u[~prob.mask])
        u = v
    assert (u == 1).sum() > 5                        # the battery is not vacuous


def test_covariate_slice_follows_time(oracle):
    # slice 0 zero, slice 1 huge: no infection at t = 0, 3 mod 2 = 1 infects
    u0 = np.zeros((12, 10), np.int8)
    u0[6, 5] = 1
    cov = np.stack([np.zeros((12, 10)), np.full((12, 10), 1e12)], axis=2)
    rule = M.FieldEpidemic(M.UniformKernel(1), covariate="c")
    prob = CAProblem(rule, u0, 4, {"beta": 1.0}, covariates={"c": cov})
    assert np.array_equal(step(oracle, prob, t0=0), u0)
    assert np.array_equal(step(oracle, prob, t0=2), u0)
    v = step(oracle, prob, t0=3)
    assert v[5, 5] == v[7, 5] == v[6, 4] == v[6, 6] == 1 and (v == 1).sum() == 5


def test_infection_rate(oracle):
    # rows alternate I / S on a torus: every S cell sees field 2/5 under the
    # radius-1 uniform kernel, so it is infected with probability 1 - exp(-0.4 beta)
    u0 = np.zeros((256, 128), np.int8)
    u0[0::2, :] = 1
    for form in ("bernoulli", "stated"):
        prob = CAProblem(M.FieldEpidemic(M.UniformKernel(1), form=form), u0, 1, {"beta": 1.5}, seed=3)
        v = step(oracle, prob)
        q = 1 - math.exp(-0.4 * 1.5)
        frac = (v[1::2, :] == 1).mean()
        n = v[1::2, :].size
        assert abs(frac - q) < 4 * math.sqrt(q * (1 - q) / n), (form, frac, q)


def test_stated_form_inverts_the_draw(oracle):
    # one cell: the stated form infects exactly when u >= 1 - q (1 - q > 0), the
    # Bernoulli form exactly when u < q: over many seeds both match those tests
    u0 = np.zeros((3, 3), np.int8)
    u0[1, 0] = 1
    hits = {"bernoulli": 0, "stated": 0}
    for seed in range(400):
        for form in hits:
            prob = CAProblem(M.FieldEpidemic(M.UniformKernel(1), form=form), u0, 1, {"beta": 2.0},
                             seed=seed, space=M.Grid(3, 3, boundary=M.Constant(0)))
            v = step(oracle, prob)
            u = oracle.uniform(seed, 5, 0, 1)  # cell (1, 1): cellid 5
            q = 1 - math.exp(-2.0 * 0.2)
            want = (u >= 1 - q) if form == "stated" else (u < q)
            assert v[1, 1] == int(want)
            hits[form] += int(want)
    assert 0 < hits["bernoulli"] < 400 and 0 < hits["stated"] < 400


def test_sir_recovery_uses_its_own_draw(oracle):
    u0 = np.ones((40, 40), np.int8)
    prob = CAProblem(M.FieldEpidemic(M.UniformKernel(1), recover=True), u0, 1,
                     {"beta": 1.0, "gamma": 0.25}, seed=4)
    v = step(oracle, prob)
    removed = (v == 2).mean()
    assert set(np.unique(v)) <= {1, 2} and abs(removed - 0.25) < 4 * math.sqrt(0.25 * 0.75 / 1600)
    # removed cells stay removed
    assert (step(oracle, prob, u=v)[v == 2] == 2).all()
    # the draw of recovery is k = 2 for the Bernoulli form
    for cell in range(1, 20):
        assert (v.ravel(order="F")[cell - 1] == 2) == (oracle.uniform(4, cell, 0, 2) < 0.25)


def test_float32_form_uses_the_24_bit_draw(oracle):
    u0 = np.zeros((64, 64), np.int8)
    u0[0::2, :] = 1
    prob = CAProblem(M.FieldEpidemic(M.UniformKernel(1), precision=32), u0, 1,
                     {"beta": np.float32(1.5)}, seed=8)
    v = step(oracle, prob)
    # decide each S cell with the Float32 formula and the 24-bit draw
    q = np.float32(1) - np.exp(-np.float32(1.5) * np.float32(0.4), dtype=np.float32)
    for y in range(0, 64, 7):
        for x in range(1, 64, 2):
            cid = x + 64 * y + 1
            assert v[x, y] == int(oracle.uniform32(8, cid, 0, 1) < q)
