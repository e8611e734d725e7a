"""Host-side logic and the C ABI's shape, without a GPU: the library loads and
exports every declared symbol, check_problem refuses what the reference refuses,
and the model mirror parses rules as the reference does."""
import ctypes
import os
import re

import numpy as np
import pytest

from cahip import _lib as L
from cahip import model as M
from cahip.backend import make_desc
from cahip.problem import CAProblem
from helpers import golden, ltl_rule

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = golden("reference_vectors.json")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cahip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cah_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES)


def test_supports():
    lib = L.load()
    assert lib.cah_supports(M.Synchronous.code) and lib.cah_supports(M.Checkerboard.code)
    assert lib.cah_supports(M.Stochastic.code)
    assert not lib.cah_supports(M.RandomSequential.code)
    assert not lib.cah_supports(M.Margolus.code)


def _check(prob, scheme=None, **kw):
    d, keep = make_desc(prob, scheme or M.Synchronous(), **kw)
    return L.load().cah_check_problem(ctypes.byref(d)), d, keep


def test_check_problem_accepts_the_configs():
    assert _check(CAProblem(M.Life(), np.zeros((512, 512), np.int8), 1))[0] == 0
    assert _check(CAProblem(ltl_rule(), np.zeros((64, 64), np.int8), 1))[0] == 0
    p = CAProblem(M.ForestFire(), np.zeros((64, 64), np.int8), 1, {"p": 0.03, "f": 1e-4},
                  space=M.Grid(64, 64, boundary=M.Reflecting()))
    assert _check(p)[0] == 0
    assert _check(CAProblem(M.Lenia(), np.zeros((64, 64), np.float32), 1))[0] == 0


def test_check_problem_refusals():
    lib = L.load()
    prob = CAProblem(M.Life(), np.zeros((8, 8), np.int8), 1)
    rc, d, keep = _check(prob)
    d.ndims = 3
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_INCOMPATIBLE_SPACE
    d.ndims = 2
    d.scheme = M.RandomSequential.code
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_UNSUPPORTED
    assert "Synchronous" in L.last_error()
    d.scheme = 0
    d.rule = 99
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_ARGUMENT
    d.rule = L.RULE_LIFE
    d.dtype = L.DTYPE_F32
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_ARGUMENT
    d.dtype = L.DTYPE_I8
    d.boundary[0] = L.BOUNDARY_CONSTANT
    d.boundary_value[0] = 300.0
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_INCOMPATIBLE_SPACE
    d.boundary_value[0] = 0.5
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_INCOMPATIBLE_SPACE
    d.boundary_value[0] = -1.0
    assert lib.cah_check_problem(ctypes.byref(d)) == 0
    d.scheme = M.Stochastic.code
    d.scheme_p = 0.0
    assert lib.cah_check_problem(ctypes.byref(d)) == L.CAH_ERR_ARGUMENT


def test_model_refusals():
    with pytest.raises(M.IncompatibleSpaceError):
        CAProblem(M.Life(), np.zeros((3, 3), np.int8), 1, space=M.Grid(3, 3, boundary=M.Constant(0.5)))
    with pytest.raises(M.InvalidStateError):
        CAProblem(M.Life(), np.full((4, 4), 2, np.int8), 1)
    with pytest.raises(M.InvalidStateError):
        CAProblem(M.Lenia(), np.full((8, 8), 2.0, np.float32), 1)
    with pytest.raises(M.StateShapeError):
        CAProblem(M.Life(), np.zeros((4, 4), np.int8), 1, space=M.Torus(4, 5))
    with pytest.raises(M.ArgumentError):
        CAProblem(M.Lenia(method="fft"), np.zeros((8, 8), np.float32), 1, space=M.Grid(8, 8))
    with pytest.raises(M.ArgumentError):
        M.Moore(0)
    with pytest.raises(M.ArgumentError):
        M.Offsets((0, 0), (1, 0))
    with pytest.raises(M.ArgumentError):
        M.Stochastic(1.5)
    with pytest.raises(M.ArgumentError):
        M.Life([64], [2])


def test_rule_strings():
    for text, want in G["rulestrings"]["cases"]:
        assert M.Life(text).rulestring() == want
    for bad in G["rulestrings"]["parse_errors"]:
        with pytest.raises(M.RuleParseError):
            M.Life(bad)
    for name, b, s, C in G["generations"]["cases"]:
        g = M.Generations(name)
        assert g.states == C and g.birth == sum(1 << x for x in b) and g.survive == sum(1 << x for x in s)
    for k, v in M.LIFE_PRESETS.items():
        assert M.Life(k).rulestring() == v
    assert M.Generations("345/2/4").rulestring() == "B2/S345/C4"


def test_desc_contents():
    prob = CAProblem(M.Lenia(), np.zeros((40, 30), np.float32), 1, seed=5)
    d, keep = make_desc(prob, M.Synchronous())
    assert d.n_taps == 516 and d.dtype == L.DTYPE_F32 and d.seed == 5
    # read offsets are the negated kernel offsets (src/environment.jl:61-62)
    offs, w = M.kernel_weights(M.LeniaKernel(13), prob.space)
    assert d.tap_offsets[0] == -offs[0][0] and d.tap_offsets[1] == -offs[0][1]
    assert abs(sum(d.tap_weights[i] for i in range(516)) - 1) < 1e-14


def test_ensemble_seeds():
    from cahip.problem import EnsembleProblem
    prob = CAProblem(M.Life(), np.zeros((4, 4), np.int8), 1, seed=G["ensemble_seeds"]["prob_seed"])
    ep = EnsembleProblem(prob)
    assert [ep.prob_func(prob, i).seed for i in range(1, 5)] == G["ensemble_seeds"]["trajectory_seeds"]
