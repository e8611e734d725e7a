"""A simulation server over the HIP backend: continuous batching of CAProblem
requests with admission by a device-memory budget (the C side is
cah_server_* in include/cahip.h).

    srv = Server(budget_bytes=200e9, chunk=64)
    rid = srv.submit(prob)            # any rule/lattice/tspan the backend runs
    while not srv.done(rid):
        srv.tick()                    # admit, queue chunks, retire; never blocks
    u = srv.fetch(rid)                # the final state, as solve() would return it

The traffic it serves is what ensembles and abc_rejection generate
(src/calibration.jl:66-91): many independent problems arriving over time.
"""
import ctypes

import numpy as np

from . import _lib as L
from . import model as M
from .backend import HIPBackend, check, make_desc, make_params


class Server:
    def __init__(self, budget_bytes=200e9, chunk=64, device=0, lenia_field="mfma_f16"):
        self.be = HIPBackend(device, lenia_field=lenia_field)
        self.lib = self.be.lib
        h = ctypes.c_void_p()
        check(self.lib.cah_server_create(self.be.ctx(), int(budget_bytes), int(chunk),
                                         ctypes.byref(h)))
        self.h = h
        self.meta = {}

    def submit(self, prob, scheme=None, u0=None):
        """Queue prob; u0, when given, replaces prob.u0 as the initial state (any
        contents of the state's element type, as set_state! takes them)."""
        scheme = M.Synchronous() if scheme is None else scheme
        d, keep = make_desc(prob, scheme, self.be.lenia_field)
        cp = make_params(prob, prob.p)
        u0 = np.asfortranarray(prob.u0 if u0 is None else np.asarray(u0, dtype=prob.u0.dtype))
        if u0.shape != tuple(prob.space.dims):
            raise M.StateShapeError(f"expected size {prob.space.dims}, got {u0.shape}")
        rid = ctypes.c_uint64()
        check(self.lib.cah_server_submit(self.h, ctypes.byref(d), u0.ctypes.data, u0.nbytes,
                                         prob.tspan[0], prob.tspan[1], ctypes.byref(cp),
                                         ctypes.byref(rid)))
        self.meta[rid.value] = (prob.space.dims, u0.dtype)
        return rid.value

    def reserve(self, device_bytes, pinned_bytes):
        """Size the server's device and pinned host memory up front
        (cah_server_reserve), as a server sized for its traffic does at start."""
        check(self.lib.cah_server_reserve(self.h, int(device_bytes), int(pinned_bytes)))

    def tune(self, chunk, tick_us):
        """Steps per request per tick at most, and the estimated device time of a
        tick in microseconds."""
        check(self.lib.cah_server_tune(self.h, int(chunk), float(tick_us)))

    def tick(self):
        a, p = ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.cah_server_tick(self.h, ctypes.byref(a), ctypes.byref(p)))
        return a.value, p.value

    def poll(self, cap=1024):
        """Ids of the requests that finished since the last poll."""
        ids = (ctypes.c_uint64 * cap)()
        n = ctypes.c_int32()
        out = []
        while True:
            check(self.lib.cah_server_poll(self.h, ids, cap, ctypes.byref(n)))
            out.extend(ids[:n.value])
            if n.value < cap:
                return out

    def status(self, rid):
        s = ctypes.c_int32()
        check(self.lib.cah_server_status(self.h, rid, ctypes.byref(s)))
        return s.value

    def done(self, rid):
        return self.status(rid) >= L.REQ_DONE

    def fetch(self, rid):
        dims, dtype = self.meta[rid]
        out = np.empty(dims, dtype=dtype, order="F")
        check(self.lib.cah_server_fetch(self.h, rid, out.ctypes.data, out.nbytes))
        return out

    def times(self, rid):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.cah_server_times(self.h, rid, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(c)))
        return a.value, b.value, c.value

    def clock(self):
        return self.lib.cah_server_clock()

    def close(self):
        if self.h:
            self.lib.cah_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
