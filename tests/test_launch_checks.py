"""The host-side launch checks (cah_check_launches, cah_check_batch_launches,
csrc/launch_check.hpp), on the CPU: every kernel launch of a problem is
replayed with fake device addresses through a model of the kernel's
addressing, which must keep every read and write inside the buffer it
belongs to, cover the lattice with the grid and fit the LDS.

Replayed here: the exact inputs of the two GPU tests held off the GPU after a
lost box (the 50 x 20 forest fire at t = 10 with p or f zero, Orbium on a 96 x
96 torus under every Lenia field method for 600 steps), the benchmark
configurations at full size, the batched configuration (256 trajectories)
under every tile the autotuner tries, and the conformance battery's rules,
schemes, edges and odd shapes.  A native self-test plants geometry mistakes in
the model's inputs and requires each to be caught."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, seed_trajectory
from helpers import forest_problem, golden, ltl_rule, rng_state
from test_oracle import CASES, SCHEMES

FIELDS = ["seq", "f32", "fft_f32", "mfma_f16x2", "mfma_f16", "mfma_bf16"]


def checked(prob, scheme=None, field="f32", tune=None, **kw):
    rep = HIPBackend(lenia_field=field, tune=tune).check_launches(prob, scheme, **kw)
    assert rep and all(n > 0 and c > 0 for n, c in rep.values()), rep
    return rep


# ------------------------------------------------------------------ the held-off GPU tests' inputs

@pytest.mark.parametrize("k", [1, 2])
def test_replay_forest_printed_draw(k):
    """test_forest_draw_matches_reference_printed_value: 50 x 20, Reflecting,
    seed 2026, tspan (10, 11), p (k = 1) or f (k = 2) just above and at the
    printed draw, the other zero."""
    u_ref = {1: 0.038246647972453474, 2: 0.403975113636703}[k]
    for q in (float(np.nextafter(u_ref, 1.0)), u_ref):
        u0 = np.full((50, 20), 0 if k == 1 else 1, np.int8)
        p, f = (q, 0.0) if k == 1 else (0.0, q)
        prob = forest_problem((50, 20), p=p, f=f, seed=2026, u0=u0, tspan=(10, 11))
        rep = checked(prob)
        assert list(rep) == ["k_forest_vn1"] and rep["k_forest_vn1"][0] == 1


@pytest.mark.parametrize("field", FIELDS)
def test_replay_orbium_masses(field):
    """test_orbium_reference_masses: Orbium at (38, 38) on a 96 x 96 torus, 600
    steps, every field method."""
    g = golden("reference_vectors.json")["orbium_masses"]
    rows = np.array(golden("reference_vectors.json")["orbium"]["rows"], dtype=np.float32)
    u = np.zeros(tuple(g["dims"]), np.float32)
    c1, c2 = (c - 1 for c in g["corner"])
    u[c1:c1 + 20, c2:c2 + 20] = rows
    prob = CAProblem(M.Lenia(), u, g["times"][-1])
    rep = checked(prob, field=field)
    assert sum(n for n, _ in rep.values()) >= 600


# ------------------------------------------------------------------ full-size configurations

def test_configs_full_size():
    z = np.zeros
    checked(CAProblem(M.Life("conway"), z((512, 512), np.int8), 1000))
    checked(CAProblem(ltl_rule(), z((8192, 8192), np.int8), 200), nsteps=2)
    checked(forest_problem((16384, 16384), tspan=500), nsteps=2)
    checked(CAProblem(M.FieldEpidemic(M.AnisotropicKernel(M.GaussianKernel(2.0), np.pi / 4, 1.0)),
                      z((2048, 2048), np.int8), 3, {"beta": 0.05}))
    for field in FIELDS:
        checked(CAProblem(M.Lenia(), z((2048, 2048), np.float32), 500), field=field, nsteps=2)


def batch_sides(n=256, seed=2026):
    u = np.random.default_rng(seed).random(n)
    return [int(round(64 * 32 ** x)) for x in u]


@pytest.mark.parametrize("tile", [(0, 0)] + [(v, r) for v in (1, 2) for r in (4, 8, 16, 32, 64)])
def test_batched_forest_config(tile):
    """256 forest fires with sides round(64 * 32^u), seeds 1..256: every tile the
    autotuner tries."""
    base = [forest_problem((s, s), tspan=2, seed=0) for s in batch_sides()]
    probs = [seed_trajectory(p, i + 1) for i, p in enumerate(base)]
    rep = HIPBackend().check_batch_launches(probs, tile=tile)
    assert set(rep) == {"k_forest_batch"}


def test_batched_life_forms():
    sides = [(33, 17), (250, 3), (1000, 41), (64, 64)]
    for rule, kernels in ((M.Life("conway"), {"k_life_tb_batch", "k_life_batch"}), (ltl_rule(), {"k_life_batch"})):
        probs = [CAProblem(rule, rng_state(i, d, [0, 1]), 3, space=M.Torus(*d)) for i, d in enumerate(sides)]
        assert set(HIPBackend().check_batch_launches(probs)) == kernels


# ------------------------------------------------------------------ the conformance battery and odd shapes

@pytest.mark.parametrize("case", [c[0] for c in CASES])
@pytest.mark.parametrize("si", range(len(SCHEMES)))
def test_rules_and_schemes(case, si):
    rule, space, vals = dict(CASES)[case]()
    u = rng_state(11 + si, space.dims, vals)
    p = {"p": 0.2, "f": 0.05} if isinstance(rule, M.ForestFire) else None
    mask = None
    if si == 1:
        mask = np.ones(space.dims, bool)
        mask.flat[::7] = False
    checked(CAProblem(rule, u, (3, 7), p, space=space, seed=3, mask=mask), SCHEMES[si])


EDGES = [(M.Periodic(), M.Periodic()), (M.Reflecting(), M.Reflecting()), (M.Constant(1), M.Constant(0)),
         (M.Periodic(), M.Constant(2)), (M.Reflecting(), M.Periodic())]
SHAPES = [(1, 1), (3, 2), (4, 4), (63, 5), (64, 7), (248, 3), (249, 9), (497, 33), (1024, 2), (2050, 40)]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("edges", range(len(EDGES)))
def test_band_walks_and_tiles_on_odd_shapes(dims, edges):
    space = M.Lattice(dims, EDGES[edges])
    u = rng_state(1, dims, [0, 1])
    for r in (1, 2, 5):
        for tune in (None, (1, 4), (2, 16), (2, 1)):
            if r == 5 and tune and tune[0] == 1:
                continue
            rule = M.Life("B3/S23", neighbourhood=M.Moore(r)) if r < 5 else ltl_rule()
            checked(CAProblem(rule, u, 3, space=space), tune=tune)
    for tune in (None, (1, 3), (2, 8)):
        checked(CAProblem(M.ForestFire(), u, 3, {"p": 0.1, "f": 0.01}, space=space, seed=2), tune=tune)
    checked(CAProblem(M.Generations("brians_brain"), u, 3, space=space))


@pytest.mark.parametrize("dims", [(20, 20), (96, 96), (100, 60), (116, 116), (120, 90), (200, 136), (333, 65)])
@pytest.mark.parametrize("edges", range(len(EDGES)))
def test_lenia_on_odd_shapes(dims, edges):
    b = EDGES[edges]
    space = M.Lattice(dims, tuple(M.Constant(0.25) if isinstance(x, M.Constant) else x for x in b))
    u = np.zeros(dims, np.float32)
    for field in FIELDS:
        if field == "fft_f32" and not all(isinstance(x, M.Periodic) for x in b):
            continue
        for R in (5, 13, 17):
            checked(CAProblem(M.Lenia(R=R), u, 2, space=space), field=field)


def test_one_dimensional_lattices():
    for b in (M.Periodic(), M.Reflecting(), M.Constant(1)):
        u = rng_state(2, (77,), [0, 1, 2])
        checked(CAProblem(M.Life("B1/S1", neighbourhood=M.Moore(2)), u % 2, 4, space=M.Lattice((77,), (b,))))
        checked(CAProblem(M.ForestFire(), u, 4, {"p": 0.2, "f": 0.1}, space=M.Lattice((77,), (b,))))


# ------------------------------------------------------------------ the checker catches planted mistakes

@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="needs hipcc")
def test_checker_catches_planted_mistakes(tmp_path):
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "cellularautomatacore.jl_amd", "csrc")
    exe = str(tmp_path / "launch_check_selftest")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-O1", "-I", csrc,
                    os.path.join(here, "native", "launch_check_selftest.hip"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
