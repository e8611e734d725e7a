"""ctypes wrapper of the CPU oracle (oracle/build/libcaref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test and the
CPU-baseline leg of bench.py, as the checker; the product never imports it.
Descriptors are passed as pointers to the same cah_desc structure the product
builds (include/cahip.h), so both sides see identical inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libcaref.so")


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "build/libcaref.so"], check=True)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        P, I32, I64, U64, D = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                               ctypes.c_double)
        sig = {
            "ref_philox4x32": (None, [P, P, P]),
            "ref_uniform": (D, [U64, I64, I64, I64]),
            "ref_uniform32": (ctypes.c_float, [U64, I64, I64, I64]),
            "ref_map_axis": (I64, [I32, I64, I64, P]),
            "ref_offsets": (I32, [I32, I32, I32, P]),
            "ref_lenia_kernel": (I32, [I32, P, I32, P, P]),
            "ref_advance_plain": (ctypes.c_int, [P, P, I64, I64, P]),
            "ref_update_cell": (ctypes.c_int, [P, P, I64, I64, I64, P]),
            "ref_advance_halo": (ctypes.c_int, [P, P, I64, I64, P, I32]),
            "ref_lenia_field": (ctypes.c_int, [P, P, P]),
            "ref_epi_field": (ctypes.c_int, [P, P, P]),
            "ref_fill_bernoulli": (None, [U64, I64, D, P]),
            "ref_fill_uniform_box": (None, [U64, I64, I64, I64, I64, I64, I64, P]),
            "ref_desc_layout": (ctypes.c_int, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def philox4x32(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    load().ref_philox4x32(c, k, o)
    return tuple(o)


def uniform(seed, cellid, t, k=1):
    return load().ref_uniform(seed, cellid, t, k)


def uniform32(seed, cellid, t, k=1):
    return load().ref_uniform32(seed, cellid, t, k)


def map_axis(kind, j, n):
    out = ctypes.c_int32()
    m = load().ref_map_axis(kind, j, n, ctypes.byref(out))
    return m, bool(out.value)


def offsets(kind, ndims, r):
    buf = (ctypes.c_int32 * (2 * (2 * r + 1) ** 2))()
    m = load().ref_offsets(kind, ndims, r, buf)
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(m)]


def lenia_kernel(R, peaks=(1.0,)):
    n = (2 * R + 1) ** 2
    pk = (ctypes.c_double * len(peaks))(*peaks)
    offs = (ctypes.c_int32 * (2 * n))()
    w = (ctypes.c_double * n)()
    m = load().ref_lenia_kernel(R, pk, len(peaks), offs, w)
    return [(offs[2 * i], offs[2 * i + 1]) for i in range(m)], np.array(w[:m])


def _buf(u):
    u = np.asfortranarray(u)
    return u, u.ctypes.data


def advance_plain(desc, u, t0, nsteps, params=None):
    """ReferenceBackend steps; returns the new state (u is not modified)."""
    out = np.array(u, order="F", copy=True)
    rc = load().ref_advance_plain(ctypes.byref(desc), out.ctypes.data, t0, nsteps,
                                  ctypes.byref(params) if params is not None else None)
    assert rc == 0
    return out


def advance_halo(desc, u, t0, nsteps, params=None, nthreads=1):
    """ArrayBackend steps (halo-padded, threaded slabs); returns the new state."""
    out = np.array(u, order="F", copy=True)
    rc = load().ref_advance_halo(ctypes.byref(desc), out.ctypes.data, t0, nsteps,
                                 ctypes.byref(params) if params is not None else None, nthreads)
    assert rc == 0
    return out


def update_cell(desc, u, i1, i2, t, params=None):
    u = np.asfortranarray(u)
    return load().ref_update_cell(ctypes.byref(desc), u.ctypes.data, i1, i2, t,
                                  ctypes.byref(params) if params is not None else None)


def lenia_field(desc, u):
    u = np.asfortranarray(u, dtype=np.float32)
    out = np.empty(u.shape, dtype=np.float64, order="F")
    load().ref_lenia_field(ctypes.byref(desc), u.ctypes.data, out.ctypes.data)
    return out


def epi_field(desc, u):
    """The field epidemic's kernel field over the Int8 state u."""
    u = np.asfortranarray(u, dtype=np.int8)
    out = np.empty(u.shape, dtype=np.float64, order="F")
    load().ref_epi_field(ctypes.byref(desc), u.ctypes.data, out.ctypes.data)
    return out


def fill_bernoulli(seed, dims, p):
    out = np.empty(dims, dtype=np.int8, order="F")
    load().ref_fill_bernoulli(seed, out.size, p, out.ctypes.data)
    return out


def fill_uniform_box(seed, dims, lo1, len1, lo2, len2):
    out = np.empty(dims, dtype=np.float32, order="F")
    n2 = dims[1] if len(dims) > 1 else 1
    load().ref_fill_uniform_box(seed, dims[0], n2, lo1, len1, lo2, len2, out.ctypes.data)
    return out


def desc_layout():
    buf = (ctypes.c_int64 * 64)()
    k = load().ref_desc_layout(buf)
    return list(buf[:k])
