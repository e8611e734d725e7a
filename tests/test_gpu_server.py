"""The serving form: requests of mixed rules, lattices and tspans arriving while
others run give, each, exactly the result of solving that problem alone, and
for the Int8 rules exactly the CPU oracle's (the threaded ArrayBackend
restatement); Lenia requests equal their single GPU run, which
test_gpu_lenia.py holds to the FP64 oracle per step."""
import numpy as np
import pytest

from cahip import _lib as L
from cahip import model as M
from cahip.backend import HIPBackend
from cahip.problem import CAProblem, final_state, solve
from cahip.server import Server
from helpers import ltl_rule, oracle_final

pytestmark = pytest.mark.gpu


def mixed_requests(n, seed=4):
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(n):
        side = int(rng.integers(40, 300))
        side2 = int(rng.integers(40, 300))
        steps = int(rng.integers(5, 60))
        kind = i % 6
        if kind == 0:
            u = rng.integers(0, 2, (side, side2)).astype(np.int8)
            probs.append(CAProblem(M.Life(), u, steps, seed=i))
        elif kind == 1:
            u = rng.choice(np.array([0, 1, 2], np.int8), (side, side2))
            probs.append(CAProblem(M.ForestFire(), u, (3, 3 + steps), {"p": 0.05, "f": 0.01},
                                   seed=100 + i, space=M.Grid(side, side2, boundary=M.Reflecting())))
        elif kind == 2:
            u = np.zeros((side, side2), np.float32)
            u[10:30, 10:30] = rng.random((20, 20)).astype(np.float32)
            probs.append(CAProblem(M.Lenia(), u, min(steps, 20), seed=i))
        elif kind == 3:
            u = (rng.random((side, side2)) < 0.4).astype(np.int8)
            probs.append(CAProblem(ltl_rule(), u, min(steps, 15), seed=i))
        elif kind == 4:  # no batched kernel: launches of its own
            u = rng.integers(0, 3, (side, side2)).astype(np.int8)
            probs.append(CAProblem(M.Generations("B2/S/C3"), u, steps, seed=i,
                                   space=M.Grid(side, side2, boundary=M.Constant(1))))
        else:  # Larger-than-Life of radius 2 with a Constant(1) edge: its own batch
            u = rng.integers(0, 2, (side, side2)).astype(np.int8)
            probs.append(CAProblem(M.Life(range(5, 8), range(4, 9), neighbourhood=M.Moore(2)), u, steps, seed=i,
                                   space=M.Grid(side, side2, boundary=M.Constant(1))))
    return probs


def assert_exact(oracle, results, ids, probs, schemes=None, lenia_field="mfma_f16"):
    """Every served result equals the problem solved alone on the GPU and, for an
    Int8 rule, the oracle's final state."""
    be = HIPBackend(lenia_field=lenia_field)
    schemes = schemes or [M.Synchronous()] * len(probs)
    for rid, p, sc in zip(ids, probs, schemes):
        alone = final_state(solve(p, sc, backend=be, save_everystep=False))
        assert np.array_equal(results[rid], alone), f"request {rid} ({p.rule}) differs from its single run"
        if not isinstance(p.rule, M.Lenia):
            assert np.array_equal(results[rid], oracle_final(oracle, p, sc)), \
                f"request {rid} ({p.rule}) differs from the oracle"


@pytest.mark.parametrize("chunk,tick_us", [(0, 0), (7, 0), (7, 400)])
def test_mix_equals_alone(oracle, chunk, tick_us):
    probs = mixed_requests(30)
    # a budget that holds only a few lattices at once, so requests queue
    srv = Server(budget_bytes=3_000_000, chunk=chunk, lenia_field="mfma_f16")
    srv.tune(chunk, tick_us)
    ids = []
    for k, p in enumerate(probs):  # requests arrive while earlier ones run
        ids.append(srv.submit(p))
        if k % 3 == 2:
            srv.tick()
    results, order = {}, []
    for _ in range(100000):
        srv.tick()
        for rid in srv.poll():
            assert srv.done(rid) and rid not in results
            results[rid] = srv.fetch(rid)
            order.append(rid)
        if len(results) == len(ids):
            break
    assert len(results) == len(ids)
    assert srv.poll() == []
    # poll reports requests in the order their ticks finished
    t_done = [srv.times(rid)[2] for rid in order]
    assert t_done == sorted(t_done)
    assert_exact(oracle, results, ids, probs)
    for rid in ids:
        t_sub, t_adm, t_done = srv.times(rid)
        assert t_sub <= t_adm <= t_done
    srv.close()


def test_fft_and_stochastic_requests(oracle):
    """Requests on the FFT field and the stochastic scheme (the loop family)."""
    rng = np.random.default_rng(9)
    probs, schemes = [], []
    for i in range(4):
        u = np.zeros((96, 80), np.float32)
        u[20:50, 20:50] = rng.random((30, 30)).astype(np.float32)
        probs.append(CAProblem(M.Lenia(method="fft"), u, 12, seed=i))
        schemes.append(M.Synchronous())
        v = rng.integers(0, 2, (70, 90)).astype(np.int8)
        probs.append(CAProblem(M.Life(), v, 25, seed=50 + i))
        schemes.append(M.Stochastic(0.5))
    srv = Server(budget_bytes=50_000_000, chunk=5, lenia_field="mfma_f16")
    ids = [srv.submit(p, sc) for p, sc in zip(probs, schemes)]
    results = {}
    for _ in range(100000):
        srv.tick()
        for rid in ids:
            if rid not in results and srv.done(rid):
                results[rid] = srv.fetch(rid)
        if len(results) == len(ids):
            break
    assert_exact(oracle, results, ids, probs, schemes)
    srv.close()


def test_budget_refuses_a_request_larger_than_it():
    srv = Server(budget_bytes=10_000, chunk=8)
    with pytest.raises(M.OutOfMemoryError):
        srv.submit(CAProblem(M.Life(), np.zeros((200, 200), np.int8), 5))
    srv.close()


def test_fetch_from_threads_while_ticking(oracle):
    """Client threads copy results out (lattices of 5 MB and more, which the host
    copy splits over its helper threads) while the serving thread submits and
    ticks; the budget holds a few lattices, so fetches free room for waiting
    requests."""
    import queue
    import threading
    rng = np.random.default_rng(21)
    probs = []
    for i in range(9):
        if i % 3 == 0:
            u = rng.integers(0, 2, (2304, 2112)).astype(np.int8)
            probs.append(CAProblem(M.Life(), u, 12, seed=i))
        elif i % 3 == 1:
            u = rng.choice(np.array([0, 1, 2], np.int8), (2200, 2400))
            probs.append(CAProblem(M.ForestFire(), u, 9, {"p": 0.05, "f": 0.001}, seed=i))
        else:
            u = np.zeros((1200, 1100), np.float32)
            u[100:400, 200:500] = rng.random((300, 300)).astype(np.float32)
            probs.append(CAProblem(M.Lenia(), u, 5, seed=i))
    srv = Server(budget_bytes=40_000_000, chunk=4, lenia_field="mfma_f16")
    work, results, errors = queue.Queue(), {}, []

    def client():
        while (rid := work.get()) is not None:
            try:
                results[rid] = srv.fetch(rid)
            except Exception as e:  # reported below
                errors.append(e)

    threads = [threading.Thread(target=client) for _ in range(3)]
    for t in threads:
        t.start()
    ids, polled = [], 0
    try:
        for p in probs:
            ids.append(srv.submit(p))
            srv.tick()
        for _ in range(1000000):
            if polled == len(ids):
                break
            srv.tick()
            for rid in srv.poll():
                work.put(rid)
                polled += 1
    finally:
        for _ in threads:
            work.put(None)
        for t in threads:
            t.join()
    assert not errors, errors
    assert polled == len(ids) and len(results) == len(ids)
    assert_exact(oracle, results, ids, probs)
    for rid in ids:
        assert srv.status(rid) == L.REQ_FETCHED
    srv.close()


@pytest.mark.parametrize("chunk", [1, 3])
def test_forest_values_outside_the_rule(oracle, chunk):
    """Forest-fire requests whose states hold values the rule never makes (3, -1,
    5), and requests whose Constant edge is 3 or 2, served over several ticks
    next to requests holding only 0, 1 and 2: the short byte tests must wait
    until a step has cleaned a lattice, and never run beside a Constant(3)
    edge.  Each result equals the oracle's."""
    from cahip.backend import make_desc, make_params
    from cahip.server import Server as S
    rng = np.random.default_rng(17)
    cases = []  # (dims, edge, values)
    for i, (dims, edge, vals) in enumerate([
            ((150, 97), M.Reflecting(), [0, 1, 2, 3, -1, 5, 1, 1]),
            ((300, 64), M.Constant(3), [0, 1, 2, 1]),
            ((97, 130), M.Constant(2), [0, 1, 1, 2]),
            ((64, 64), M.Reflecting(), [0, 1, 1, 2]),
            ((200, 33), M.Constant(3), [0, 1, 2, 3, -1, 5])]):
        cases.append((dims, edge, rng.choice(np.array(vals, np.int8), dims)))
    srv = S(budget_bytes=50_000_000, chunk=chunk)
    ids, refs = [], []
    for i, (dims, edge, u) in enumerate(cases):
        space = M.Grid(*dims, boundary=edge)
        prob = CAProblem(M.ForestFire(), np.zeros(dims, np.int8), (2, 12), {"p": 0.2, "f": 0.02}, space=space,
                         seed=30 + i)
        ids.append(srv.submit(prob, u0=u))
        d, keep = make_desc(prob, M.Synchronous())
        refs.append(oracle.advance_halo(d, np.asfortranarray(u), 2, 10, make_params(prob, prob.p), 4))
        srv.tick()
    results = {}
    for _ in range(100000):
        srv.tick()
        for rid in srv.poll():
            results[rid] = srv.fetch(rid)
        if len(results) == len(ids):
            break
    for rid, ref in zip(ids, refs):
        assert np.array_equal(results[rid], ref), f"request {rid} differs from the oracle"
    srv.close()
