"""Benchmark of the synchronous sweep on one MI355X.

Workload (one "step"): one synchronous step of each of
  - Larger-than-Life Moore(5) B34..45/S34..58, Torus 8192 x 8192, Int8,
    Bernoulli(0.5) initial state from seed 1;
  - the forest fire, VonNeumann(1), Reflecting 16384 x 16384, Int8, all empty,
    p = 0.03, f = 1e-4, Philox seed 7;
both lattices resident in HBM.  `value` = cell updates per second over both, in
GCUPS.  Kernel durations come from HIP events recorded on each state's own
stream around its launches.  Each lattice first runs `--warmup` steps and then
at least WARM_MS of further untimed steps (the device's clocks settle over
tens of milliseconds of load), then WINDOWS timed windows of exactly `--steps`
steps each; the figures are the median window's, and the spread over the
windows is reported.  The CPU baseline is the oracle's restatement of the
reference's threaded ArrayBackend, timed on a few steps of each lattice.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cellularautomatacore.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP16_PEAK_TFLOPS = 2500.0  # dense FP16 / BF16 MFMA
WARM_MS = 200.0            # untimed device work per lattice before the timed windows
WINDOWS = 5                # timed windows of --steps steps each; the median is reported
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, in T lane-instructions/s
VALU_PEAK_T = 256 * 4 * 16 * 2.4e9 / 1e12
# k_forest_vn1<1, true, 0, true> (gfx950 ISA): 181 VALU instructions in the walk's row body of 4
# cells (4 draws, the rule, the store; the branch taken on a tie of the draw's high word left out)
FOREST_VALU_PER_DRAW = 181 / 4
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_ltl8192_forest16384.json")
LENIA_TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_lenia_mfma_2048.json")
COPY_N = 16384 * 16384  # bytes of the calibration copy (one 16384^2 Int8 lattice)
KERNELS = {"ltl": "k_life_swar<5, 2, true, 1>", "forest": "k_forest_vn1<1, true, 0, true>",
           "copy": "k_copy_rows<16>"}
ALGO_BYTES = {"ltl": 2 * 8192 * 8192, "forest": 2 * 16384 * 16384, "copy": 2 * COPY_N}


def copy_calibration(be, st, reps=3):
    """A plain device copy of known size that scales the HBM counters: the forest
    lattice (16384^2 Int8, dense rows) copied to a device buffer by k_copy_rows<16>,
    COPY_N bytes read and COPY_N written per launch."""
    import torch
    from cahip.backend import check
    dst = torch.empty(COPY_N, dtype=torch.int8, device="cuda")
    for _ in range(reps):
        check(be.lib.cah_state_get_device(st.handle, ctypes.c_void_p(dst.data_ptr()), COPY_N))
    torch.cuda.synchronize()


def pmc_summary(fetch_csv, write_csv, cmd):
    """HBM bytes per launch of the swept kernels from two rocprofv3 --pmc runs of
    bench.py --copy-calibration (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch),
    each scaled by what the counter reads for the calibration copy of known size.
    Writes TRAFFIC_JSON and a Markdown table beside it."""
    import csv

    def means(path, counter):
        acc = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                for key, name in KERNELS.items():
                    if name in r["Kernel_Name"]:
                        acc.setdefault(key, []).append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}

    fetch, write = means(fetch_csv, "FETCH_SIZE"), means(write_csv, "WRITE_SIZE")
    sf = COPY_N / 1024 / fetch["copy"][0]
    sw = COPY_N / 1024 / write["copy"][0]
    out = {"command": cmd, "scale_fetch": sf, "scale_write": sw}
    rows = ["| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes/launch (scaled) "
            "| algorithmic bytes | ratio |", "|---|---|---|---|---|---|---|"]
    for key in ("ltl", "forest", "copy"):
        f, nf = fetch[key]
        w, nw = write[key]
        b = (f * sf + w * sw) * 1024
        out[key] = {"kernel": KERNELS[key], "fetch_kib": f, "write_kib": w,
                    "bytes_per_launch": round(b), "algorithmic_bytes": ALGO_BYTES[key]}
        rows.append(f"| {KERNELS[key]} | {nf}/{nw} | {f:.0f} | {w:.0f} | {b:.4g} | "
                    f"{ALGO_BYTES[key]:.4g} | {b / ALGO_BYTES[key]:.3f} |")
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(out, fh, indent=1)
    md = TRAFFIC_JSON[:-5] + ".md"
    with open(md, "w") as fh:
        fh.write("# HBM traffic per launch (rocprofv3 --pmc)\n\n"
                 f"Commands (one counter per run, then this summary):\n\n```\n{cmd}\n```\n\n"
                 f"Scale factors from the calibration copy ({COPY_N} bytes read and written "
                 f"per launch): FETCH x{sf:.4f}, WRITE x{sw:.4f}.\n\n" + "\n".join(rows) + "\n")
    return out


def trace_summary(trace_csv, out_md, cmd, top=30):
    """Kernel durations per (kernel, grid size) from a rocprofv3 --kernel-trace
    CSV (one row per dispatch), so that the benchmark's lattices, which share
    kernels with the other configurations, are told apart by their launch
    size: calls, mean, median, min and max in microseconds, and the share of
    the traced kernel time.  Writes a Markdown table to out_md."""
    import csv
    groups, total = {}, 0.0
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            dims = lambda k: (int(r.get(k + "_X") or 1) * int(r.get(k + "_Y") or 1) *  # noqa: E731
                              int(r.get(k + "_Z") or 1)) if k + "_X" in r else int(r.get(k) or 0)
            key = (r["Kernel_Name"], dims("Grid_Size"), dims("Workgroup_Size"))
            groups.setdefault(key, []).append(d)
            total += d
    rows = sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:top]
    lines = ["| Kernel | Grid (work-items) | Block | Calls | Mean (us) | Median (us) | Min (us) | Max (us) | Share |",
             "|---|---|---|---|---|---|---|---|---|"]
    for (name, grid, wg), ds in rows:
        ds = sorted(ds)
        lines.append(f"| `{name[:90]}` | {grid} | {wg} | {len(ds)} | {sum(ds) / len(ds):.2f} | "
                     f"{ds[len(ds) // 2]:.2f} | {ds[0]:.2f} | {ds[-1]:.2f} | {100 * sum(ds) / total:.1f}% |")
    with open(out_md, "w") as fh:
        fh.write("# Kernel durations by launch size (rocprofv3 --kernel-trace)\n\n"
                 f"Command:\n\n```\n{cmd}\n```\n\n" + "\n".join(lines) + "\n")


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def make_states(be):
    from cahip import model as M
    from cahip.backend import check
    from cahip.problem import CAProblem
    ltl = CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)),
                    np.zeros((8192, 8192), np.int8), 200, space=M.Torus(8192, 8192), seed=1)
    forest = CAProblem(M.ForestFire(), np.zeros((16384, 16384), np.int8), 500,
                       {"p": 0.03, "f": 1e-4}, space=M.Grid(16384, 16384, boundary=M.Reflecting()),
                       seed=7)
    s_ltl = be.allocate_state(ltl, M.Synchronous())
    check(be.lib.cah_state_fill_bernoulli(s_ltl.handle, 1, 0.5))
    s_for = be.allocate_state(forest, M.Synchronous())
    return (ltl, s_ltl), (forest, s_for)


def host_cores():
    """The host cores this job may use: the CPUs of its affinity mask, capped by
    a cgroup CPU quota and by OMP_NUM_THREADS when they are set (a multi-GPU
    host grants each GPU's job a share of its cores and says so there).
    Returns (cores, description)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    n = aff
    if quota:
        n = min(n, quota)
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n, {"cpus_visible": os.cpu_count(), "affinity": aff, "cgroup_cpu_quota": quota,
               "omp_num_threads": omp or None, "threads_used": n}


def cpu_baseline(nthreads, steps=2):
    """The oracle's ArrayBackend restatement on the host cores, a few steps of each lattice."""
    from cahip import model as M
    from cahip.backend import make_desc, make_params
    from cahip.problem import CAProblem
    from oracle import caref
    caref.load()
    u = caref.fill_bernoulli(1, (8192, 8192), 0.5)
    ltl = CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)), u, 1,
                    space=M.Torus(8192, 8192), seed=1)
    d1, k1 = make_desc(ltl, M.Synchronous())
    t = time.perf_counter()
    caref.advance_halo(d1, u, 0, steps, None, nthreads)
    t_ltl = (time.perf_counter() - t) / steps
    del u, ltl
    z = np.zeros((16384, 16384), np.int8, order="F")
    fo = CAProblem(M.ForestFire(), z, 1, {"p": 0.03, "f": 1e-4},
                   space=M.Grid(16384, 16384, boundary=M.Reflecting()), seed=7)
    d2, k2 = make_desc(fo, M.Synchronous())
    t = time.perf_counter()
    caref.advance_halo(d2, z, 0, steps, make_params(fo, fo.p), nthreads)
    t_for = (time.perf_counter() - t) / steps
    cells = 8192 * 8192 + 16384 * 16384
    return {"value": cells / (t_ltl + t_for) / 1e9, "unit": "GCUPS", "cores": nthreads,
            "kind": "port",
            "sample": f"{steps} steps of each lattice (8192^2 LtL, 16384^2 forest fire) with the "
                      "oracle's C restatement of ArrayBackend(threaded=true): halo-padded, serial "
                      "halo fill, min(n_last, 4*threads) slabs",
            "ltl_ms_per_step": t_ltl * 1e3, "forest_ms_per_step": t_for * 1e3}


def timed(ev, stream, fn):
    """Milliseconds between HIP events recorded on `stream` around fn()."""
    a, b = ev.event(), ev.event()
    ev.sync()
    ev.record(a, stream)
    fn()
    ev.record(b, stream)
    ev.sync()
    return ev.elapsed_ms(a, b)


def copy_peak(ev, nbytes, reps=40):
    """The device's copy rate at a sweep launch's working set: torch's copy_ (a
    blit kernel) of nbytes to another buffer, `reps` times between HIP events on
    torch's stream; bytes read + written per second, in GB/s."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(5):
        dst.copy_(src)

    def run():
        for _ in range(reps):
            dst.copy_(src)

    ms = timed(ev, stream, run)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes * reps / (ms * 1e-3) / 1e9


def bench_life512(be, ev, steps=1000):
    """Conway Life B3/S23, Torus 512^2, Bernoulli(0.5) seed 1, 1000 steps."""
    from cahip import model as M
    from cahip.backend import check, make_params
    from cahip.problem import CAProblem
    prob = CAProblem(M.Life("conway"), np.zeros((512, 512), np.int8), steps,
                     space=M.Torus(512, 512), seed=1)
    st = be.allocate_state(prob, M.Synchronous())
    check(be.lib.cah_state_fill_bernoulli(st.handle, 1, 0.5))
    cp = make_params(prob, None)
    run = lambda: check(be.lib.cah_advance(st.handle, 0, steps, ctypes.byref(cp)))  # noqa: E731
    run()
    ms = timed(ev, be.stream(st), run)
    be.free_state(st)
    cells = 512 * 512
    return {"workload": "Conway Life Moore(1) Torus 512^2 Bernoulli(0.5) seed 1, 1000 steps",
            "us_per_step": round(ms * 1e3 / steps, 3),
            "gcups": round(cells * steps / (ms * 1e-3) / 1e9, 2),
            "kernel": "k_life_tb: 16 steps a launch in LDS, one bit a cell",
            "bound": "launch latency and per-step barriers (0.5 MiB working set stays in L2)"}


def batch_sides(n=256, seed=2026):
    """Square sides log-uniform in [64, 2048]: round(64 * 32^u), u from numpy's
    default_rng(2026) (the batched workload's documented generator)."""
    u = np.random.default_rng(seed).random(n)
    return [int(round(64 * 32 ** x)) for x in u]


def bench_batched(be, ev, steps=300, n=256):
    """256 forest-fire trajectories (seeds 1..256, Reflecting, all empty,
    p=0.03, f=1e-4), sides log-uniform in [64, 2048], 300 steps: one launch per
    step for all of them under a captured hipGraph, the tile (cells a lane, rows a
    wave) chosen by the batch autotuner (cah_batch_autotune), against the plain
    per-trajectory loop."""
    from cahip import _lib as L
    from cahip import model as M
    from cahip.backend import check, make_desc, make_params
    from cahip.problem import CAProblem, seed_trajectory
    sides = batch_sides(n)
    base = [CAProblem(M.ForestFire(), np.zeros((s, s), np.int8), steps, {"p": 0.03, "f": 1e-4},
                      space=M.Grid(s, s, boundary=M.Reflecting()), seed=0) for s in sides]
    probs = [seed_trajectory(p, i + 1) for i, p in enumerate(base)]
    cells = sum(s * s for s in sides)
    descs, keeps = zip(*[make_desc(p, M.Synchronous()) for p in probs])
    arr = (L.CahDesc * n)(*descs)
    cp = make_params(probs[0], probs[0].p)
    h = ctypes.c_void_p()
    check(be.lib.cah_batch_create(be.ctx(), n, arr, ctypes.byref(h)))
    s = ctypes.c_void_p()
    check(be.lib.cah_batch_stream(h, ctypes.byref(s)))
    for i, p in enumerate(probs):
        check(be.lib.cah_batch_set(h, i, p.u0.ctypes.data, p.u0.nbytes))
    # the tile (cells a lane, rows a wave) chosen by timing 20 steps of each candidate
    # (the trial steps are undone)
    tuned_vb, tuned = ctypes.c_int32(), ctypes.c_int32()
    check(be.lib.cah_batch_autotune(h, 0, 20, ctypes.byref(cp), ctypes.byref(tuned_vb), ctypes.byref(tuned)))
    check(be.lib.cah_batch_advance(h, 0, 2, ctypes.byref(cp), 1))  # warm-up (graph path)
    check(be.lib.cah_batch_sync(h))
    for i, p in enumerate(probs):  # restart every trajectory from its empty lattice
        check(be.lib.cah_batch_set(h, i, p.u0.ctypes.data, p.u0.nbytes))
    ms_batch = timed(ev, s.value, lambda: check(
        be.lib.cah_batch_advance(h, 0, steps, ctypes.byref(cp), 1)))
    batched_final = []
    for i, p in enumerate(probs):
        out = np.empty(p.space.dims, np.int8, order="F")
        check(be.lib.cah_batch_get(h, i, out.ctypes.data, out.nbytes))
        batched_final.append(out)
    be.lib.cah_batch_destroy(h)

    def plain(sync_every_step):
        finals = []
        ev.sync()
        t0 = time.perf_counter()
        for p in probs:
            st = be.allocate_state(p, M.Synchronous())
            if sync_every_step:  # advance! then synchronize, as the KA GPUBackend does
                for t in range(steps):
                    check(be.lib.cah_advance(st.handle, t, 1, ctypes.byref(cp)))
                    be.sync(st)
            else:
                check(be.lib.cah_advance(st.handle, 0, steps, ctypes.byref(cp)))
            finals.append(be.current_state(st))
            be.free_state(st)
        return time.perf_counter() - t0, finals
    t_step_sync, finals_a = plain(True)
    t_queued, finals_b = plain(False)
    same = all(np.array_equal(a, b) for a, b in zip(batched_final, finals_a)) and \
        all(np.array_equal(a, b) for a, b in zip(batched_final, finals_b))
    return {"workload": f"{n} forest-fire trajectories, seeds 1..{n}, sides round(64*32^u) "
                        f"(u ~ default_rng(2026)), Reflecting, all empty, {steps} steps",
            "total_cells": cells,
            "batched_ms": round(ms_batch, 3),
            "batched_sims_per_s": round(n / (ms_batch * 1e-3), 1),
            "batched_gcups": round(cells * steps / (ms_batch * 1e-3) / 1e9, 2),
            "plain_step_sync_sims_per_s": round(n / t_step_sync, 1),
            "plain_queued_sims_per_s": round(n / t_queued, 1),
            "speedup_vs_plain_step_sync": round(t_step_sync / (ms_batch * 1e-3), 2),
            "speedup_vs_plain_queued": round(t_queued / (ms_batch * 1e-3), 2),
            "autotuned_tile": {"cells_a_lane": 4 * tuned_vb.value, "rows_a_wave": tuned.value},
            "identical_trajectories": bool(same)}


def bench_field_epidemic(be, ev, steps=100):
    """The field epidemic on a 2048^2 torus: SI, AnisotropicKernel(GaussianKernel(2),
    pi/4, 1) (113 taps), 0.1% of cells infected from the build's generator (seed 3),
    beta 0.05; ms/step (field + rule) for Float64 and Float32 fields, HIP events on
    the state's stream."""
    from cahip import model as M
    from cahip.backend import check, make_params
    from cahip.problem import CAProblem
    n = 2048
    out = {"workload": f"field epidemic SI, Torus {n}^2, AnisotropicKernel(GaussianKernel(2), pi/4, 1), "
                       f"0.1% infected (generator seed 3), beta 0.05, {steps} steps"}
    for precision, beta in ((64, 0.05), (32, np.float32(0.05))):
        rule = M.FieldEpidemic(M.AnisotropicKernel(M.GaussianKernel(2.0), math.pi / 4, 1.0),
                               precision=precision)
        prob = CAProblem(rule, np.zeros((n, n), np.int8), steps, {"beta": beta}, seed=5)
        cp = make_params(prob, prob.p)
        st = be.allocate_state(prob, M.Synchronous())
        check(be.lib.cah_state_fill_bernoulli(st.handle, 3, 0.001))
        check(be.lib.cah_advance(st.handle, 0, 3, ctypes.byref(cp)))
        ms = timed(ev, be.stream(st), lambda: check(
            be.lib.cah_advance(st.handle, 3, steps, ctypes.byref(cp))))
        infected = be.state_counts(st).get(1, 0)
        be.free_state(st)
        out[f"f{precision}"] = {"ms_per_step": round(ms / steps, 4), "taps": len(prob.taps[0]),
                                "gcups": round(n * n * steps / (ms * 1e-3) / 1e9, 2),
                                "infected_after": infected}
    return out


def bench_counts(be, ev, reps=20):
    """cah_state_counts on the 16384^2 forest fire after 20 steps: ms per call
    (kernels + 2 KiB back), and the state's bytes read per second."""
    from cahip import model as M
    from cahip.backend import check, make_params
    from cahip.problem import CAProblem
    n = 16384
    prob = CAProblem(M.ForestFire(), np.zeros((n, n), np.int8), 20, {"p": 0.03, "f": 1e-4},
                     space=M.Grid(n, n, boundary=M.Reflecting()), seed=7)
    st = be.allocate_state(prob, M.Synchronous())
    check(be.lib.cah_advance(st.handle, 0, 20, ctypes.byref(make_params(prob, prob.p))))
    for _ in range(3):
        counts = be.state_counts(st)
    ms = timed(ev, be.stream(st), lambda: [be.state_counts(st) for _ in range(reps)]) / reps
    be.free_state(st)
    return {"workload": f"state_counts of a {n}^2 Int8 forest fire after 20 steps", "ms_per_call": round(ms, 4),
            "gbs_state_read": round(n * n / (ms * 1e-3) / 1e9, 1), "counts": counts}


def bench_sparse(be, ev, steps=200):
    """The sparse sweep against the dense one on invasions of a 16384^2 lattice
    (the SparseBackend's use, vignettes/10_backends/10_backends.md:175-197): a fire
    burning out from the centre of a full forest (forest fire, background 1, p 0.01,
    f 0, Reflecting, seed 7) and an R-pentomino on a torus (Conway Life, background
    0); tiles of 32 x 32.  ms/step over `steps` steps from the seed state, the tiles
    stored at the end, and whether both sweeps end in the same state."""
    from cahip import model as M
    from cahip.backend import HIPSparseBackend, check, make_params
    from cahip.problem import CAProblem
    n = 16384
    fire = np.ones((n, n), np.int8)
    fire[n // 2, n // 2] = 2
    life = np.zeros((n, n), np.int8)
    life[n // 2:n // 2 + 3, n // 2:n // 2 + 3] = np.array([[0, 1, 1], [1, 1, 0], [0, 1, 0]], np.int8)
    cases = {"forest_fire": (CAProblem(M.ForestFire(), fire, steps, {"p": 0.01, "f": 0.0}, seed=7,
                                       space=M.Grid(n, n, boundary=M.Reflecting())), 1),
             "life_r_pentomino": (CAProblem(M.Life("conway"), life, steps, space=M.Torus(n, n)), 0)}
    out = {"workload": f"{n}^2: forest fire burning out of a full forest (p 0.01, f 0), and an R-pentomino "
                       f"under Conway Life on a torus; {steps} steps, tiles 32 x 32"}
    del fire, life
    for name, (prob, bg) in cases.items():
        cp = make_params(prob, prob.p)
        r = {}
        finals = {}
        for kind, b in (("dense", be), ("sparse", HIPSparseBackend(bg, tilesize=32))):
            st = b.allocate_state(prob, M.Synchronous())
            ms = timed(ev, b.stream(st), lambda: check(b.lib.cah_advance(st.handle, 0, steps, ctypes.byref(cp))))
            r[f"{kind}_ms_per_step"] = round(ms / steps, 4)
            if kind == "sparse":
                r["stored_tiles"] = b.stored_tiles(st)
                r["tiles"] = (n // 32) ** 2
            finals[kind] = b.state_counts(st)
            b.free_state(st)
        r["speedup"] = round(r["dense_ms_per_step"] / r["sparse_ms_per_step"], 1)
        r["same_counts"] = finals["dense"] == finals["sparse"]
        out[name] = r
    return out


def bench_lenia(be, ev, steps=500):
    """Lenia R=13 (mu 0.15, sigma 0.015, dt 0.1), Float32 Torus 2048^2, uniform[0,1]
    in the central 256^2 patch: ms/step per field method; (i) the per-step max-abs
    difference from the FP64 sequential field (the path pinned to the oracle): one
    step of the method against one FP64 step from the same state, at every one of
    the 500 steps of the method's own trajectory (compared on the device); (ii) the
    trajectory divergence at t = 500 (max-abs and mass against the FP64 run)."""
    import torch
    from cahip import model as M
    from cahip.backend import HIPBackend, check, make_params
    from cahip.problem import CAProblem
    n = 2048
    prob = CAProblem(M.Lenia(), np.zeros((n, n), np.float32), steps, space=M.Torus(n, n), seed=1)
    cp = make_params(prob, None)
    ref_be = HIPBackend(0, lenia_field="seq")
    ref = ref_be.allocate_state(prob, M.Synchronous())
    nbytes = n * n * 4
    cur, mine, theirs = (torch.empty(n * n, dtype=torch.float32, device="cuda") for _ in range(3))
    out, seq_final = {}, None
    for field in ("seq", "mfma_f16", "mfma_f16x2", "mfma_bf16", "fft_f32", "f32"):
        fb = HIPBackend(0, lenia_field=field)
        st = fb.allocate_state(prob, M.Synchronous())
        check(fb.lib.cah_state_fill_uniform_box(st.handle, 1, 896, 256, 896, 256))
        init = fb.current_state(st)
        err = torch.zeros((), dtype=torch.float64, device="cuda")
        for t in range(steps):  # the trajectory, one step at a time
            if field != "seq":
                torch.cuda.synchronize()  # the last comparison has read the buffers
                check(fb.lib.cah_state_get_device(st.handle, ctypes.c_void_p(cur.data_ptr()), nbytes))
                check(ref_be.lib.cah_state_set_device(ref.handle, ctypes.c_void_p(cur.data_ptr()), nbytes))
                check(ref_be.lib.cah_advance(ref.handle, t, 1, ctypes.byref(cp)))
            check(fb.lib.cah_advance(st.handle, t, 1, ctypes.byref(cp)))
            if field != "seq":
                check(fb.lib.cah_state_get_device(st.handle, ctypes.c_void_p(mine.data_ptr()), nbytes))
                check(ref_be.lib.cah_state_get_device(ref.handle, ctypes.c_void_p(theirs.data_ptr()), nbytes))
                err = torch.maximum(err, (mine.double() - theirs.double()).abs().max())
        final = fb.current_state(st).astype(np.float64)
        fb.set_state(st, init)
        ms = timed(ev, fb.stream(st), lambda: check(
            fb.lib.cah_advance(st.handle, 0, steps, ctypes.byref(cp))))
        fb.free_state(st)
        if field == "seq":
            seq_final = final
        out[field] = {"ms_per_step": round(ms / steps, 4),
                      "gcups": round(n * n * steps / (ms * 1e-3) / 1e9, 2),
                      "per_step_max_abs_vs_fp64": float(err.item()),
                      "final_mass": float(final.sum()),
                      "final_max_abs_vs_fp64": float(np.abs(final - seq_final).max()),
                      "final_mass_rel_diff_vs_fp64": abs(float(final.sum()) - float(seq_final.sum())) /
                      float(seq_final.sum())}
        taps = len(prob.taps[0])
        sec = ms / steps * 1e-3
        algo_tflops = 2 * taps * n * n / sec / 1e12
        out[field]["hbm_gbs_algorithmic"] = round(8 * n * n / sec / 1e9, 1)
        if field.startswith("mfma"):
            # one 64 x 64 block: 4 waves x (row-offset pairs over the rows holding
            # weights) x 12 16x16x32 MFMAs (6 input chunks, 4 strips, 3 chunk offsets
            # each; 16384 flop an MFMA), three per product when split
            dys = np.array([o[1] for o, w in zip(*prob.taps) if w != 0])
            npairs = (int(dys.max() - dys.min()) + 2) // 2
            per_block = 4 * npairs * 12 * 2 * 16 * 16 * 32 * (3 if field == "mfma_f16x2" else 1)
            exec_tflops = per_block * (n // 64) ** 2 / sec / 1e12
            traffic = None
            if os.path.exists(LENIA_TRAFFIC_JSON):
                with open(LENIA_TRAFFIC_JSON) as f:
                    traffic = json.load(f).get(field, {}).get("bytes_per_launch")
            out[field]["roofline"] = {
                "bound": "mfma", "kernel": "k_lenia_mfma", "achieved": round(algo_tflops, 1),
                "peak": FP16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(algo_tflops / FP16_PEAK_TFLOPS, 4),
                "algorithmic_flop_per_cell": 2 * taps,
                "executed_tflops": round(exec_tflops, 1),
                "executed_frac": round(exec_tflops / FP16_PEAK_TFLOPS, 4),
                "traffic": traffic,
                "note": "achieved counts the direct sum's 2 x 516 flop a cell; executed counts every MFMA "
                        "the implicit GEMM issues, the Toeplitz band's zeros included; traffic: HBM bytes a "
                        "launch from FETCH_SIZE / WRITE_SIZE (profiles/hbm_traffic_lenia_mfma_2048.json)"}
        else:
            out[field]["tflops_algorithmic"] = round(algo_tflops, 2)
    ref_be.free_state(ref)
    return {"workload": "Lenia R=13 Float32 Torus 2048^2, central 256^2 uniform patch, 500 steps",
            "methods": out}


SERVING_MIX = ("Life B3/S23 Torus Bernoulli(0.5); forest fire VN(1) Reflecting all-empty "
               "p=0.03 f=1e-4; Lenia R=13 Torus (MFMA FP16) uniform central side/8 patch{ltl}; "
               "in turn; sides round({lo} * {r}^u) (log-uniform in [{lo}, {hi}]) square; "
               "tspan (0, U{{50..{tmax}}})")


def serving_problems(n, seed=5, lo=64, hi=1024, tmax=300, ltl=True):
    """n requests of the mix SERVING_MIX, in turn.  With lo=128, hi=4096, tmax=500 and
    no Larger-than-Life requests it is the serving configuration's stream; with the
    defaults it is the traffic ensembles and ABC rejection generate (many
    independent small problems).  numpy default_rng(seed)."""
    from cahip import model as M
    from cahip.problem import CAProblem
    rng = np.random.default_rng(seed)
    rule_ltl = M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5))
    kinds = (0, 1, 2, 3) if ltl else (0, 1, 3)
    out = []
    for i in range(n):
        side = int(round(lo * (hi / lo) ** rng.random()))
        steps = int(rng.integers(50, tmax + 1))
        kind = kinds[i % len(kinds)]
        # initial states in the column-major order the backend takes, as a
        # Julia caller's arrays are (no transposing copy when submitted), and
        # written, not lazily mapped zero pages (np.full, not np.zeros), as a
        # request's data arriving over a connection is
        if kind == 0:
            u = np.asfortranarray(rng.integers(0, 2, (side, side), dtype=np.int8))
            p = CAProblem(M.Life("conway"), u, steps, seed=i)
        elif kind == 1:
            p = CAProblem(M.ForestFire(), np.full((side, side), 0, np.int8, order="F"), steps,
                          {"p": 0.03, "f": 1e-4}, seed=1000 + i,
                          space=M.Grid(side, side, boundary=M.Reflecting()))
        elif kind == 2:
            u = np.asfortranarray(rng.integers(0, 2, (side, side), dtype=np.int8))
            p = CAProblem(rule_ltl, u, steps, seed=i)
        else:
            u = np.full((side, side), 0, np.float32, order="F")
            k = max(side // 8, 1)
            a = (side - k) // 2
            u[a:a + k, a:a + k] = rng.random((k, k), dtype=np.float32)
            p = CAProblem(M.Lenia(), u, steps, seed=i)
        out.append(p)
    return out


def poisson_arrivals(n, rate, seed=6):
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n))


def serve(probs, arrivals, chunk=32, budget=200e9, tick_us=None, clients=4, inline_bytes=0, submitters=4):
    """The server (cah_server_*) fed in arrival order; per request the time its
    result reached host memory, and the results.  As a server's connection
    threads would, `submitters` threads hand the requests over, each taking the
    next one in arrival order at its arrival time (cah_server_submit copies its
    initial state into the server's pinned memory without holding the server's
    lock), and `clients` threads copy out results of `inline_bytes` and more
    (by default every result), while the serving thread ticks and polls.  A
    result the serving thread fetched itself would hold back the next tick: with
    the results below 4 MB fetched inline the device sat idle for about half of
    the stream's last 80 ms (1,020 against 1,086 sims/s on the 400 requests)."""
    import concurrent.futures as cf
    import threading
    from cahip.server import Server
    srv = Server(budget_bytes=budget, chunk=chunk, lenia_field="mfma_f16")
    if tick_us:
        srv.tune(chunk, tick_us)
    # sized for the stream at start, as a server sized for its traffic is: the
    # lattices (two buffers each, rows padded) and the pinned staging of initial
    # states and results
    state_bytes = sum(p.u0.nbytes for p in probs)
    srv.reserve(min(budget, 2.5 * state_bytes), 2.2 * state_bytes)
    for p in probs[:4]:  # warm-up: one request of each kind
        rid = srv.submit(p)
        while not srv.done(rid):
            srv.tick()
        srv.fetch(rid)
    srv.poll()
    n = len(probs)
    fin, res, index = [0.0] * n, [None] * n, {}
    left = n
    pool = cf.ThreadPoolExecutor(max_workers=clients) if clients else None
    futures, unknown = [], []

    def take(rid, j):
        res[j] = srv.fetch(rid)
        fin[j] = time.perf_counter() - start

    nxt, lock = [0], threading.Lock()

    def submitter():
        while True:
            with lock:
                j = nxt[0]
                nxt[0] += 1
            if j >= n:
                return
            wait = arrivals[j] - (time.perf_counter() - start)
            if wait > 0:
                time.sleep(wait)
            index[srv.submit(probs[j])] = j

    start = time.perf_counter()
    subs = [threading.Thread(target=submitter) for _ in range(submitters)]
    for t in subs:
        t.start()
    while left:
        srv.tick()
        done, unknown = unknown + srv.poll(), []
        for rid in done:
            j = index.get(rid)
            if j is None:  # finished before its submitter recorded it
                unknown.append(rid)
                continue
            if pool and probs[j].u0.nbytes >= inline_bytes:
                futures.append(pool.submit(take, rid, j))
            else:
                take(rid, j)
            left -= 1
    for t in subs:
        t.join()
    for f in futures:
        f.result()
    if pool:
        pool.shutdown()
    srv.close()
    return np.array(fin), res


def plain(probs, arrivals):
    """The plain per-problem GPU sweep: one request at a time in arrival order,
    each allocated, advanced, read back and freed (what solve() does)."""
    from cahip import model as M
    from cahip.backend import HIPBackend, check, make_params
    be = HIPBackend(0, lenia_field="mfma_f16")

    def one(p):
        st = be.allocate_state(p, M.Synchronous())
        cp = make_params(p, p.p)
        check(be.lib.cah_advance(st.handle, p.tspan[0], p.tspan[1] - p.tspan[0], ctypes.byref(cp)))
        u = be.current_state(st)
        be.free_state(st)
        return u

    for p in probs[:4]:
        one(p)
    fin, res = [], []
    start = time.perf_counter()
    for ta, p in zip(arrivals, probs):
        while time.perf_counter() - start < ta:
            pass
        res.append(one(p))
        fin.append(time.perf_counter() - start)
    return np.array(fin), res


def latency_pair(probs, arr, budget=200e9):
    """Served and plain runs of one arrival stream: p50/p99 of arrival -> result in
    host memory, sims/s to the last result, and whether the results agree."""
    fs, rs = serve(probs, arr, budget=budget)
    fp, rp = plain(probs, arr)
    ls, lp = (fs - arr) * 1e3, (fp - arr) * 1e3
    n = len(probs)
    return {"served": {"p50_ms": round(float(np.percentile(ls, 50)), 2),
                       "p99_ms": round(float(np.percentile(ls, 99)), 2),
                       "sims_per_s": round(n / fs.max(), 1)},
            "plain": {"p50_ms": round(float(np.percentile(lp, 50)), 2),
                      "p99_ms": round(float(np.percentile(lp, 99)), 2),
                      "sims_per_s": round(n / fp.max(), 1)},
            "identical_results": bool(all(np.array_equal(a, b) for a, b in zip(rs, rp)))}


def bench_gated(be, ev, steps=50):
    """The sweep beyond the two headline kernels, each lattice timed alone (a
    window of `steps` steps after 100 ms of untimed ones): Generations, the
    Stochastic and Checkerboard gates and a mask on the gated band walks
    (k_life_gated, k_forest_gated), and a neighbourhood no band walk takes
    (VonNeumann(2), and Moore(10) with 440 neighbours) on the window-sum kernel
    (k_box_i8).  Roofline:
    algorithmic bytes (one Int8 read and write a cell, and a mask byte read
    where there is a mask) against 8 TB/s."""
    from cahip import model as M
    from cahip.backend import check, make_params
    from cahip.problem import CAProblem
    n1, n2 = 8192, 16384
    torus = M.Torus(n1, n1)
    refl = M.Grid(n2, n2, boundary=M.Reflecting())
    ff = {"p": 0.03, "f": 1e-4}
    i16 = np.arange(n2, dtype=np.int16)
    mask = ((i16[:, None] + i16[None, :]) % 7 != 0)  # 6 in 7 cells open, in diagonal stripes
    cases = [
        ("brians_brain_8192", "Generations brians_brain (B2/S/3) Moore(1) Torus 8192^2, Bernoulli(0.3) seed 1, "
         "Synchronous", CAProblem(M.Generations("brians_brain"), np.zeros((n1, n1), np.int8), 1, space=torus),
         M.Synchronous(), (1, 0.3), 0),
        ("life_stochastic_8192", "Life B3/S23 Moore(1) Torus 8192^2, Bernoulli(0.5) seed 1, Stochastic(0.5) seed 3",
         CAProblem(M.Life(), np.zeros((n1, n1), np.int8), 1, space=torus, seed=3), M.Stochastic(0.5), (1, 0.5), 0),
        ("ltl_checkerboard_8192", "Larger-than-Life Moore(5) B34..45/S34..58 Torus 8192^2, Bernoulli(0.5) seed 1, "
         "Checkerboard (two half-sweeps a step)",
         CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(5)), np.zeros((n1, n1), np.int8), 1,
                   space=torus), M.Checkerboard(), (1, 0.5), 0),
        ("forest_stochastic_16384", "forest fire VN(1) Reflecting 16384^2 all empty p=0.03 f=1e-4 seed 7, "
         "Stochastic(0.5)", CAProblem(M.ForestFire(), np.zeros((n2, n2), np.int8), 1, ff, space=refl, seed=7),
         M.Stochastic(0.5), None, 0),
        ("forest_masked_16384", "forest fire VN(1) Reflecting 16384^2 all empty p=0.03 f=1e-4 seed 7, "
         "Synchronous, mask (i + j) % 7 != 0",
         CAProblem(M.ForestFire(), np.zeros((n2, n2), np.int8), 1, ff, space=refl, seed=7, mask=mask),
         M.Synchronous(), None, 1),
        ("life_vn2_8192", "Life B3/S23 VonNeumann(2) Torus 8192^2, Bernoulli(0.5) seed 1, Synchronous "
         "(the window-sum kernel)", CAProblem(M.Life("B3/S23", neighbourhood=M.VonNeumann(2)),
                                               np.zeros((n1, n1), np.int8), 1, space=torus),
         M.Synchronous(), (1, 0.5), 0),
        ("ltl_moore10_8192", "Larger-than-Life Moore(10) B34..45/S34..58 Torus 8192^2, Bernoulli(0.05) seed 1, "
         "Synchronous (the window-sum kernel)",
         CAProblem(M.Life(range(34, 46), range(34, 59), neighbourhood=M.Moore(10)), np.zeros((n1, n1), np.int8), 1,
                   space=torus), M.Synchronous(), (1, 0.05), 0),
    ]
    out = {}
    for name, what, prob, scheme, fill, mask_bytes in cases:
        st = be.allocate_state(prob, scheme)
        if fill:
            check(be.lib.cah_state_fill_bernoulli(st.handle, *fill))
        kernels = sorted(be.check_launches(prob, scheme))
        cp = make_params(prob, prob.p)
        stream, t = be.stream(st), 0
        spent = 0.0
        while spent < 100.0:
            spent += timed(ev, stream, lambda: check(be.lib.cah_advance(st.handle, t, 10, ctypes.byref(cp))))
            t += 10
        ms = timed(ev, stream, lambda: check(be.lib.cah_advance(st.handle, t, steps, ctypes.byref(cp)))) / steps
        be.free_state(st)
        cells = prob.u0.size
        gbs = cells * (2 + mask_bytes) / (ms * 1e-3) / 1e9
        out[name] = {"workload": what, "kernels": kernels, "ms_per_step": round(ms, 5),
                     "gcups": round(cells / (ms * 1e-3) / 1e9, 2),
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                                  "bytes_per_cell": 2 + mask_bytes}}
    del mask
    return out


def bench_serving():
    """The serving configuration: a Poisson stream at 200 req/s (400 requests) of
    Life, forest-fire and Lenia requests in turn, sides log-uniform in [128, 4096],
    tspan U[50, 500], served under a 200 GB budget with 32-step ticks, against the
    plain per-problem sweep taking the same stream."""
    n, rate = 400, 200.0
    probs = serving_problems(n, seed=5, lo=128, hi=4096, tmax=500, ltl=False)
    arr = poisson_arrivals(n, rate, seed=6)
    out = {"workload": SERVING_MIX.format(ltl="", lo=128, r=32, hi=4096, tmax=500) +
           f"; Poisson {rate:g} req/s, {n} requests; budget 200 GB",
           "requests": n,
           "cell_updates": int(sum(p.u0.size * (p.tspan[1] - p.tspan[0]) for p in probs))}
    out.update(latency_pair(probs, arr))
    # capacity: the same requests all present at t = 0, served and plain
    cap = latency_pair(probs, np.zeros(n))
    out["capacity"] = {"served_sims_per_s": cap["served"]["sims_per_s"],
                       "plain_sims_per_s": cap["plain"]["sims_per_s"],
                       "ratio": round(cap["served"]["sims_per_s"] / cap["plain"]["sims_per_s"], 2),
                       "served_p50_ms": cap["served"]["p50_ms"], "served_p99_ms": cap["served"]["p99_ms"],
                       "identical_results": cap["identical_results"]}
    return out


def bench_serving_ensembles():
    """Ensemble-style traffic (many small problems, Larger-than-Life added):
    capacity with 400 requests present at t = 0, and latency under Poisson
    arrivals at a light load both paths sustain and at a heavy one."""
    probs = serving_problems(400)
    zero = np.zeros(len(probs))
    cap = latency_pair(probs, zero)
    out = {"workload": SERVING_MIX.format(ltl="; Larger-than-Life Moore(5) B34..45/S34..58 Torus "
                                              "Bernoulli(0.5)", lo=64, r=16, hi=1024, tmax=300),
           "capacity": {"requests": len(probs),
                        "cell_updates": int(sum(p.u0.size * (p.tspan[1] - p.tspan[0]) for p in probs)),
                        "served_sims_per_s": cap["served"]["sims_per_s"],
                        "plain_sims_per_s": cap["plain"]["sims_per_s"],
                        "ratio": round(cap["served"]["sims_per_s"] / cap["plain"]["sims_per_s"], 2),
                        "identical_results": cap["identical_results"]}}
    for name, rate, n, seed in (("light", 150.0, 300, 7), ("heavy", 1000.0, 1000, 8)):
        probs = serving_problems(n, seed=seed)
        arr = poisson_arrivals(n, rate, seed)
        out[name] = {"rate_per_s": rate, "requests": n}
        out[name].update(latency_pair(probs, arr))
    return out


PROFILE_WORKLOADS = ("ltl", "forest", "lenia_mfma_f16", "lenia_mfma_f16x2", "lenia_mfma_bf16", "lenia_fft_f32",
                     "brians_brain", "forest_stochastic", "copy", "serving_capacity", "plain_capacity")


def profile_workload(name, steps):
    """One workload alone, for profiler passes: `steps` steps (after 3 untimed
    ones) of Larger-than-Life Moore(5) 8192^2 (ltl), the forest fire 16384^2
    (forest), Lenia R=13 2048^2 with a field method (lenia_*), Brian's Brain
    Moore(1) Torus 8192^2 (brians_brain), a Stochastic(0.5) forest fire
    16384^2 (forest_stochastic), or the calibration copy (copy); prints the mean
    ms per step from HIP events.  serving_capacity / plain_capacity: the serving
    configuration's 400 requests, all present at t = 0, through the server or the
    plain per-request sweep; prints sims/s."""
    from cahip import model as M
    from cahip.backend import HIPBackend, check, make_params
    from cahip.problem import CAProblem
    if name in ("serving_capacity", "plain_capacity"):
        # the serving configuration's 400 requests, all present at t = 0
        probs = serving_problems(400, seed=5, lo=128, hi=4096, tmax=500, ltl=False)
        fin, _ = (serve if name == "serving_capacity" else plain)(probs, np.zeros(len(probs)))
        print(json.dumps({"workload": name, "requests": len(probs), "sims_per_s": round(len(probs) / fin.max(), 1)}),
              flush=True)
        return
    ev = HipEvents()
    if name in ("ltl", "forest", "copy"):
        be = HIPBackend(0)
        (ltl, s_ltl), (forest, s_for) = make_states(be)
        if name == "copy":
            copy_calibration(be, s_for, reps=steps)
            return
        prob, st, scheme = (ltl, s_ltl, M.Synchronous()) if name == "ltl" else (forest, s_for, M.Synchronous())
    else:
        field = name[len("lenia_"):] if name.startswith("lenia_") else "f32"
        be = HIPBackend(0, lenia_field=field)
        scheme = M.Synchronous()
        if name.startswith("lenia_"):
            prob = CAProblem(M.Lenia(), np.zeros((2048, 2048), np.float32), steps, space=M.Torus(2048, 2048))
            st = be.allocate_state(prob, scheme)
            check(be.lib.cah_state_fill_uniform_box(st.handle, 1, 896, 256, 896, 256))
        elif name == "brians_brain":
            prob = CAProblem(M.Generations("brians_brain"), np.zeros((8192, 8192), np.int8), steps,
                             space=M.Torus(8192, 8192))
            st = be.allocate_state(prob, scheme)
            check(be.lib.cah_state_fill_bernoulli(st.handle, 1, 0.3))
        else:
            scheme = M.Stochastic(0.5)
            prob = CAProblem(M.ForestFire(), np.zeros((16384, 16384), np.int8), steps, {"p": 0.03, "f": 1e-4},
                             space=M.Grid(16384, 16384, boundary=M.Reflecting()), seed=7)
            st = be.allocate_state(prob, scheme)
    cp = make_params(prob, prob.p)
    check(be.lib.cah_advance(st.handle, 0, 3, ctypes.byref(cp)))
    ms = timed(ev, be.stream(st), lambda: check(be.lib.cah_advance(st.handle, 3, steps, ctypes.byref(cp))))
    print(json.dumps({"workload": name, "steps": steps, "ms_per_step": round(ms / steps, 5)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (Life 512^2, batched, Lenia, serving)")
    ap.add_argument("--configs", default=None,
                    help="comma-separated subset of the other configurations to run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the plain copy that scales the HBM counters")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "COMMAND"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) and exit")
    ap.add_argument("--trace-summary", nargs=3, metavar=("TRACE_CSV", "OUT_MD", "COMMAND"),
                    help="summarise a rocprofv3 --kernel-trace CSV by kernel and launch size and exit")
    ap.add_argument("--profile", choices=PROFILE_WORKLOADS,
                    help="run one workload alone for --steps steps (for profiler passes) and exit")
    args = ap.parse_args()
    if args.profile:
        profile_workload(args.profile, args.steps)
        return
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary)))
        return
    if args.trace_summary:
        trace_summary(*args.trace_summary)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the sweep on a single MI355X "
              "(the multi-GPU slab decomposition is design only)", file=sys.stderr)
        sys.exit(2)

    from cahip.backend import HIPBackend, check, make_params
    be = HIPBackend(0)
    (ltl, s_ltl), (forest, s_for) = make_states(be)
    ev = HipEvents()
    from cahip import model as M
    sync = M.Synchronous()
    pf = make_params(forest, forest.p)
    pl = make_params(ltl, None)

    def run(st, params, t0, n):
        check(be.lib.cah_advance(st.handle, t0, n, ctypes.byref(params)))

    K, W = args.steps, args.warmup

    def reset(st, kind):
        """The lattice's initial state again: Bernoulli(0.5) from seed 1 (LtL), all empty (forest)."""
        check(be.lib.cah_state_fill_bernoulli(st.handle, *((1, 0.5) if kind == "ltl" else (7, 0.0))))

    def measure(st, params, kind):
        """At least WARM_MS of untimed steps (the clocks settle under load), then
        WINDOWS windows, each from the initial state: W untimed steps, then
        exactly K steps between HIP events on the state's own stream.  Per
        window: (device ms per step, host wall ms per step)."""
        stream, spent, t = be.stream(st), 0.0, 0
        while spent < WARM_MS:
            spent += timed(ev, stream, lambda: run(st, params, t, 20))
            t += 20
        out = []
        for _ in range(WINDOWS):
            reset(st, kind)
            run(st, params, 0, W)
            a, b = ev.event(), ev.event()
            ev.sync()
            t_wall = time.perf_counter()
            ev.record(a, stream)
            run(st, params, W, K)
            ev.record(b, stream)
            ev.sync()
            wall = time.perf_counter() - t_wall
            out.append((ev.elapsed_ms(a, b) / K, wall * 1e3 / K))
        return out

    # each lattice timed on its own (the other idle)
    win_ltl = measure(s_ltl, pl, "ltl")
    win_for = measure(s_for, pf, "forest")
    med = lambda xs: float(np.median(xs))  # noqa: E731
    ms_ltl, ms_for = med([w[0] for w in win_ltl]), med([w[0] for w in win_for])
    wall_ms = med([w[1] for w in win_ltl]) + med([w[1] for w in win_for])

    c_ltl, c_for = 8192 * 8192, 16384 * 16384
    ms_step = ms_ltl + ms_for
    value = (c_ltl + c_for) / (ms_step * 1e-3) / 1e9
    # dominant kernel: the forest fire sweep (one launch per step); algorithmic
    # bytes = one Int8 read and one Int8 write per cell
    achieved = ALGO_BYTES["forest"] / (ms_for * 1e-3) / 1e9
    valu_t = c_for * FOREST_VALU_PER_DRAW / (ms_for * 1e-3) / 1e12
    achieved_ltl = ALGO_BYTES["ltl"] / (ms_ltl * 1e-3) / 1e9
    traffic = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            traffic = json.load(f)
    if args.copy_calibration:
        copy_calibration(be, s_for)
    # measured copy rate at each kernel's working set (half its algorithmic
    # bytes read, half written), beside the nominal 8 TB/s
    cp_ltl = copy_peak(ev, ALGO_BYTES["ltl"] // 2)
    cp_for = copy_peak(ev, ALGO_BYTES["forest"] // 2)
    # the same K steps with the states handed over in host memory: one
    # cah_state_set before and one cah_state_get after each lattice's run
    host_ms = 0.0
    for st in (s_ltl, s_for):
        u = be.current_state(st)
        t = time.perf_counter()
        be.set_state(st, u)
        be.current_state(st)
        host_ms += (time.perf_counter() - t) * 1e3
    line = {
        "metric": "GCUPS + % HBM roofline (synchronous sweep)",
        "value": round(value, 3), "unit": "GCUPS", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round(wall_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8", "data": "synthetic",
        "config": {"workload": "LtL Moore(5) B34..45/S34..58 Torus 8192^2 Bernoulli(0.5) seed 1 "
                               "+ forest fire VN(1) Reflecting 16384^2 all-empty p=0.03 f=1e-4 seed 7",
                   "global_batch": 2, "seq_len": None, "parallelism": "none (1 GPU)"},
        "workloads": {
            "ltl_8192": {"gcups": round(c_ltl / ms_ltl / 1e6, 3), "ms_per_step": round(ms_ltl, 5),
                         "hbm_gbs_algorithmic": round(2 * c_ltl / ms_ltl / 1e6, 1)},
            "forest_16384": {"gcups": round(c_for / ms_for / 1e6, 3),
                             "ms_per_step": round(ms_for, 5),
                             "hbm_gbs_algorithmic": round(2 * c_for / ms_for / 1e6, 1)},
        },
        "timing": {"warm_ms_per_lattice": WARM_MS, "windows": WINDOWS,
                   "note": f"each window: the initial state, {W} untimed steps, then {K} timed steps; "
                           "the median window is reported",
                   "ltl_ms_per_step": [round(w[0], 5) for w in win_ltl],
                   "forest_ms_per_step": [round(w[0], 5) for w in win_for],
                   "spread": round(max(max(w[0] for w in ws) / min(w[0] for w in ws) - 1
                                       for ws in (win_ltl, win_for)), 4)},
        "pcie_inclusive": {"host_copy_ms": round(host_ms, 2),
                           "gcups": round((c_ltl + c_for) * K / (ms_step * K + host_ms) / 1e6, 3),
                           "note": f"{K} steps of each lattice with the initial state set from and "
                                   "the final state read back to pageable host memory"},
        "roofline": {"bound": "valu", "kernel": KERNELS["forest"], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("forest", {}).get("bytes_per_launch"),
                     "copy_gbs": round(cp_for, 1), "frac_of_copy": round(achieved / cp_for, 4),
                     "valu": {"instructions_per_draw": FOREST_VALU_PER_DRAW,
                              "achieved_t_lane_instr_s": round(valu_t, 2), "peak": round(VALU_PEAK_T, 2),
                              "frac": round(valu_t / VALU_PEAK_T, 4)},
                     "note": "achieved/peak/frac: the HBM view (2 B per cell update against 8 TB/s); the "
                             "kernel is bound by the VALU (valu: 45.25 lane-instructions per cell for "
                             "one Philox4x32-10 draw, 17 v_mad_u64_u32 / v_mul_*_u32 and 17-18 "
                             "v_bitop3_b32, the rule and the store, against the issue peak)"},
        "roofline_ltl": {"bound": "hbm", "kernel": KERNELS["ltl"], "achieved": round(achieved_ltl, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved_ltl / HBM_PEAK_GBS, 4),
                         "traffic": traffic.get("ltl", {}).get("bytes_per_launch"),
                         "copy_gbs": round(cp_ltl, 1), "frac_of_copy": round(achieved_ltl / cp_ltl, 4)},
    }
    be.free_state(s_ltl)
    be.free_state(s_for)
    if not args.no_configs:
        configs = {"life_512": lambda: bench_life512(be, ev),
                   "batched_forest": lambda: bench_batched(be, ev),
                   "lenia_2048": lambda: bench_lenia(be, ev),
                   "field_epidemic_2048": lambda: bench_field_epidemic(be, ev),
                   "counts_16384": lambda: bench_counts(be, ev),
                   "sparse_16384": lambda: bench_sparse(be, ev),
                   "gated_sweeps": lambda: bench_gated(be, ev),
                   "serving": bench_serving,
                   "serving_ensembles": bench_serving_ensembles}
        names = list(configs) if args.configs is None else args.configs.split(",")
        unknown = [k for k in names if k not in configs]
        if unknown:
            sys.exit(f"unknown configs {unknown}; known: {list(configs)}")
        line["configs"] = {k: configs[k]() for k in names}
    if not args.no_cpu:
        n, host = host_cores()
        line["cpu_baseline"] = cpu_baseline(n)
        line["cpu_baseline"]["host"] = host
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
