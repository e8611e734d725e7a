"""The batched, served and Lenia configurations at their stated sizes,
anchored to the CPU oracle.

- Batched forest fires: 256 trajectories, seeds 1..256, sides round(64 * 32^u)
  (u from default_rng(2026)), Reflecting, all empty, p = 0.03, f = 1e-4, 300
  steps, one batched launch per step: every final state against the SHA-256
  digest of the oracle's own run (tests/golden/full_size.json, written by
  tests/golden/make_full_size_goldens.py), with the default tile and with the
  autotuned one.
- Served requests: the serving configuration's stream (Life, forest fire and
  MFMA-FP16 Lenia in turn, sides log-uniform in [128, 4096], tspan U[50, 500],
  the generator bench.py uses), submitted while the server ticks: each result
  equals its single run, and each Int8 result of a side up to 640 the oracle's.
- Lenia R = 13 on a 2048^2 torus from a uniform central 256^2 patch: each step
  of the FP16 MFMA, split-FP16 MFMA and FFT-FP32 fields against one step of
  the FP64 oracle from the same state, within the stated tolerances (1e-3,
  1e-5, 1e-5 max-abs).  The oracle steps the window that holds every non-zero
  cell and 2H cells around it (the field vanishes beyond H of a non-zero
  cell, so the window's Constant(0) edge reads what the torus holds there);
  every cell outside the window must stay 0, as it does in the oracle.
"""
import hashlib

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend, check
from cahip.problem import CAProblem, final_state, seed_trajectory, solve
from cahip.server import Server
from helpers import desc_for, forest_problem, golden, oracle_final

pytestmark = pytest.mark.gpu


def digest(u):
    return hashlib.sha256(np.asfortranarray(u).tobytes(order="F")).hexdigest()


@pytest.mark.parametrize("autotune", [False, True])
def test_batched_forest_256_trajectories(autotune):
    want = golden("full_size.json")["batched_forest_256"]["final"]
    base = [forest_problem((e["side"], e["side"]), p=0.03, f=1e-4, seed=0, tspan=300) for e in want]
    probs = [seed_trajectory(p, i + 1) for i, p in enumerate(base)]
    be = HIPBackend()
    outs = be.solve_batch(probs, M.Synchronous(), autotune=autotune)
    for k, (e, u) in enumerate(zip(want, outs)):
        assert int(u.astype(np.int64).sum()) == e["sum"], f"trajectory {k + 1}: cell sum differs"
        assert digest(u) == e["sha256"], f"trajectory {k + 1} (side {e['side']}) differs from the oracle"


def test_serving_stream():
    import bench
    probs = bench.serving_problems(45, seed=5, lo=128, hi=4096, tmax=500, ltl=False)
    srv = Server(budget_bytes=200e9, chunk=32, lenia_field="mfma_f16")
    ids, results = [], {}
    for p in probs:  # one arrival per tick: requests join while others run
        ids.append(srv.submit(p))
        srv.tick()
        for rid in srv.poll():
            results[rid] = srv.fetch(rid)
    for _ in range(1000000):
        if len(results) == len(ids):
            break
        srv.tick()
        for rid in srv.poll():
            results[rid] = srv.fetch(rid)
    srv.close()
    assert len(results) == len(ids)
    be = HIPBackend(lenia_field="mfma_f16")
    return_oracle = []
    for rid, p in zip(ids, probs):
        alone = final_state(solve(p, backend=be, save_everystep=False))
        assert np.array_equal(results[rid], alone), f"request {rid} ({p.rule}, {p.space.dims}) differs"
        if not isinstance(p.rule, M.Lenia) and max(p.space.dims) <= 640:
            return_oracle.append((rid, p))
    from oracle import caref
    caref.load()
    assert return_oracle, "the stream holds small Int8 requests"
    for rid, p in return_oracle:
        assert np.array_equal(results[rid], oracle_final(caref, p)), f"request {rid} differs from the oracle"


@pytest.mark.parametrize("field,tol", [("mfma_f16", 1e-3), ("mfma_f16x2", 1e-5), ("fft_f32", 1e-5)])
def test_lenia_2048_steps_against_the_oracle(oracle, field, tol):
    n, H, steps = 2048, 12, 3
    prob = CAProblem(M.Lenia(), np.zeros((n, n), np.float32), steps, space=M.Torus(n, n), seed=1)
    be = HIPBackend(lenia_field=field)
    st = be.allocate_state(prob, M.Synchronous())
    try:
        check(be.lib.cah_state_fill_uniform_box(st.handle, 1, 896, 256, 896, 256))
        u = be.current_state(st)
        assert (u[896:1152, 896:1152] > 0).mean() > 0.99 and u.sum() == u[896:1152, 896:1152].sum()
        for t in range(steps):
            be.advance(st, prob, M.Synchronous(), None, t, 1)
            v = be.current_state(st)
            nz = np.argwhere(u != 0)
            lo, hi = nz.min(0) - 2 * H, nz.max(0) + 2 * H + 1
            assert (lo >= 0).all() and (hi <= n).all()
            win = (slice(lo[0], hi[0]), slice(lo[1], hi[1]))
            crop = np.asfortranarray(u[win])
            cp = CAProblem(M.Lenia(), crop, 1, space=M.Grid(*crop.shape, boundary=M.Constant(0.0)))
            d, keep = desc_for(cp, field_method=0)
            ref = oracle.advance_plain(d, crop, t, 1)
            err = float(np.abs(v[win].astype(np.float64) - ref).max())
            assert err < tol, f"{field}: step {t}: max-abs {err:.3g}"
            outside = v.copy()
            outside[win] = 0
            assert not outside.any(), f"{field}: step {t}: a cell outside the window changed"
            u = v
    finally:
        be.free_state(st)
