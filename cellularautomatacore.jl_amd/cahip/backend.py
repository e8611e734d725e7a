"""HIPBackend: the MI355X backend behind the reference's backend interface.

Each method is the counterpart of one generic function of the backend contract
(src/backends.jl:1-14, :89-148) and calls the C ABI of libcahip.so
(include/cahip.h).  Julia's bang functions keep their names without the bang:
advance! -> advance, set_state! -> set_state, write_cell! -> write_cell.
"""
import ctypes

import numpy as np

from . import _lib as L
from . import measures as MS
from . import model as M

_ERRORS = {
    L.CAH_ERR_ARGUMENT: M.ArgumentError,
    L.CAH_ERR_INCOMPATIBLE_SPACE: M.IncompatibleSpaceError,
    L.CAH_ERR_UNSUPPORTED: M.UnsupportedCombinationError,
    L.CAH_ERR_OUT_OF_MEMORY: M.OutOfMemoryError,
    L.CAH_ERR_STATE_SHAPE: M.StateShapeError,
    L.CAH_ERR_INVALID_STATE: M.InvalidStateError,
    L.CAH_ERR_DEVICE: M.DeviceError,
    L.CAH_ERR_BUSY: M.OutOfMemoryError,
}


def check(rc):
    if rc != L.CAH_OK:
        raise _ERRORS.get(rc, M.DeviceError)(L.last_error())
    return rc


FIELD_METHODS = {"seq": L.FIELD_DIRECT_SEQ, "f32": L.FIELD_DIRECT_F32,
                 "mfma_f16": L.FIELD_MFMA_F16, "mfma_bf16": L.FIELD_MFMA_BF16,
                 "mfma_f16x2": L.FIELD_MFMA_F16X2, "fft_f32": L.FIELD_FFT_F32}


def _p_value(p, name, what="the forest fire", default=None):
    if p is None:
        if default is not None:
            return default
        raise M.ArgumentError(f"{what} reads p.{name}, and the problem has no parameters")
    if isinstance(p, dict):
        if name not in p and default is not None:
            return default
        return p[name]
    return getattr(p, name, default) if default is not None else getattr(p, name)


def make_params(prob, p):
    """cah_params from the integrator's p: the forest fire's (p, f), the field
    epidemic's (beta, gamma)."""
    cp = L.CahParams()
    if isinstance(prob.rule, M.ForestFire):
        pv, fv = _p_value(p, "p"), _p_value(p, "f")
    elif isinstance(prob.rule, M.FieldEpidemic):
        pv = _p_value(p, "beta", "the field epidemic")
        fv = _p_value(p, "gamma", "the field epidemic") if prob.rule.recover else 0.0
    else:
        return cp
    cp.ff_p, cp.ff_f = float(pv), float(fv)
    cp.ff_p_is_f32 = int(isinstance(pv, np.float32))
    cp.ff_f_is_f32 = int(isinstance(fv, np.float32))
    return cp


def make_desc(prob, scheme, field_method=L.FIELD_DIRECT_F32):
    """A cah_desc for (prob, scheme) and the host arrays it points into."""
    d = L.CahDesc()
    keep = []
    sp = prob.space
    d.ndims = sp.ndims
    rule = prob.rule
    d.dtype = L.DTYPE_F32 if isinstance(rule, M.Lenia) else L.DTYPE_I8
    d.dims[0] = sp.dims[0]
    d.dims[1] = sp.dims[1] if sp.ndims == 2 else 1
    for a, b in enumerate(sp.boundaries):
        d.boundary[a] = b.code
        d.boundary_value[a] = float(b.value) if isinstance(b, M.Constant) else 0.0
    if sp.ndims == 1:
        d.boundary[1] = L.BOUNDARY_PERIODIC
    d.scheme = getattr(scheme, "code", -1)
    if isinstance(scheme, M.Stochastic):
        d.scheme_p = float(scheme.p)
        d.scheme_p_is_f32 = int(scheme.is_f32)
    d.seed = prob.seed
    d.rule = rule.code
    if isinstance(rule, (M.Life, M.Generations)):
        d.birth, d.survive = rule.birth, rule.survive
    if isinstance(rule, M.Generations):
        d.gen_states = rule.states
    offs = np.array([(o + (0,))[:2] for o in prob.neighbours.offsets], dtype=np.int32).ravel()
    keep.append(offs)
    d.n_offsets = len(prob.neighbours.offsets)
    d.offsets = offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    if isinstance(rule, M.Lenia):
        taps, w = prob.taps
        to = np.array(taps, dtype=np.int32).ravel()
        tw = np.ascontiguousarray(w, dtype=np.float64)
        keep += [to, tw]
        d.n_taps = len(taps)
        d.tap_offsets = to.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.tap_weights = tw.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.mu, d.sigma, d.dt = rule.mu, rule.sigma, rule.dt
        d.lenia_precision = rule.precision
        d.field_method = L.FIELD_FFT_F32 if rule.method == "fft" else field_method
    if isinstance(rule, M.FieldEpidemic):
        taps, w = prob.taps
        to = np.array(taps, dtype=np.int32).ravel()
        tw = np.ascontiguousarray(w, dtype=np.float64)
        keep += [to, tw]
        d.n_taps = len(taps)
        d.tap_offsets = to.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.tap_weights = tw.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.lenia_precision = rule.precision
        d.epi_form = L.EPI_STATED if rule.form == "stated" else L.EPI_BERNOULLI
        d.epi_recover = int(rule.recover)
        if rule.covariate is not None:
            cv = prob.covariates[rule.covariate]
            n = int(np.prod(sp.dims))
            slices = 1 if cv.ndim == sp.ndims else cv.shape[-1]
            flat = np.ascontiguousarray(cv.reshape((n, slices), order="F").T).ravel()  # slice-major
            keep.append(flat)
            d.covariate = flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            d.covariate_slices = slices
    if prob.mask is not None:
        m = np.asfortranarray(prob.mask).ravel(order="F").astype(np.uint8)
        keep.append(m)
        d.mask = m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    return d, keep


def _parse_report(text):
    """{kernel: (launches, checks)} from a launch-check report."""
    out = {}
    for line in text.splitlines():
        name, rest = line.rsplit(": ", 1)
        launches, checks = (int(x.split()[0]) for x in rest.split(", "))
        out[name] = (launches, checks)
    return out


def count_dict(counts):
    """{Int8 value: count} over the values present, from counts by byte value."""
    return {(v if v < 128 else v - 256): int(c) for v, c in enumerate(counts) if c}


class HIPState:
    """The backend state: a handle to device buffers owned by libcahip."""

    def __init__(self, handle, dims, dtype):
        self.handle, self.dims, self.dtype = handle, dims, dtype

    @property
    def nbytes(self):
        return int(np.prod(self.dims)) * np.dtype(self.dtype).itemsize


class HIPBackend:
    """HIPBackend(device=0; lenia_field="f32").

    `lenia_field` chooses how the Lenia potential is computed on the device:
    "seq" sums the taps in order in the rule's precision (as _field_at does),
    "f32" uses FP32 FMAs over a dense weight table.
    """

    def __init__(self, device=0, lenia_field="f32", tune=None):
        self.lib = L.load()
        self.device = device
        if lenia_field not in FIELD_METHODS:
            raise M.ArgumentError(f"unknown lenia_field {lenia_field!r}")
        self.lenia_field = FIELD_METHODS[lenia_field]
        self.tune = tune
        self._ctx = None

    def ctx(self):
        if self._ctx is None:
            h = ctypes.c_void_p()
            check(self.lib.cah_open(self.device, ctypes.byref(h)))
            self._ctx = h
        return self._ctx

    # supports(backend, scheme) (src/backends.jl:95, :153)
    def supports(self, scheme):
        return bool(self.lib.cah_supports(getattr(scheme, "code", -1)))

    # check_problem(backend, prob, scheme) (src/backends.jl:105)
    def check_problem(self, prob, scheme):
        d, _keep = make_desc(prob, scheme, self.lenia_field)
        check(self.lib.cah_check_problem(ctypes.byref(d)))

    # allocate_state(backend, prob, scheme) (src/backends.jl:119)
    def allocate_state(self, prob, scheme):
        d, _keep = make_desc(prob, scheme, self.lenia_field)
        h = ctypes.c_void_p()
        check(self.lib.cah_state_create(self.ctx(), ctypes.byref(d), ctypes.byref(h)))
        st = HIPState(h, prob.space.dims, prob.u0.dtype)
        if self.tune:
            check(self.lib.cah_state_tune(h, *self.tune))
        self.set_state(st, prob.u0)
        return st

    # advance!(backend, st, prob, scheme, p, t) (src/backends.jl:130), nsteps times
    def advance(self, st, prob, scheme, p, t, nsteps=1):
        cp = make_params(prob, p)
        check(self.lib.cah_advance(st.handle, int(t), int(nsteps), ctypes.byref(cp)))

    # current_state(backend, st) (src/backends.jl:139): a copy the caller may keep
    def current_state(self, st):
        out = np.empty(st.dims, dtype=st.dtype, order="F")
        check(self.lib.cah_state_get(st.handle, out.ctypes.data, st.nbytes))
        return out

    # set_state!(backend, st, u) (src/backends.jl:148)
    def set_state(self, st, u):
        u = np.asfortranarray(np.asarray(u, dtype=st.dtype))
        if tuple(u.shape) != tuple(st.dims):
            raise M.StateShapeError(f"expected size {st.dims}, got {u.shape}")
        check(self.lib.cah_state_set(st.handle, u.ctypes.data, st.nbytes))

    # read_cell / write_cell! (src/agents.jl:165, :174), 0-based index tuple
    def read_cell(self, st, I):
        v = np.zeros(1, dtype=st.dtype)
        i2 = I[1] if len(I) > 1 else 0
        check(self.lib.cah_state_read_cell(st.handle, int(I[0]), int(i2), v.ctypes.data))
        return v[0]

    def write_cell(self, st, I, value):
        v = np.array([value], dtype=st.dtype)
        i2 = I[1] if len(I) > 1 else 0
        check(self.lib.cah_state_write_cell(st.handle, int(I[0]), int(i2), v.ctypes.data))

    def sync(self, st):
        check(self.lib.cah_state_sync(st.handle))

    def stream(self, st):
        s = ctypes.c_void_p()
        check(self.lib.cah_state_stream(st.handle, ctypes.byref(s)))
        return s.value

    def free_state(self, st):
        if st.handle:
            self.lib.cah_state_destroy(st.handle)
            st.handle = None

    # -------------------------------------------------------------- summaries

    def _counts(self, st, other=None):
        """cah_state_counts: counts by byte value (256) and, with `other`, the cells
        where the two states agree."""
        counts = np.zeros(256, np.int64)
        eq = np.zeros(1, np.int64)
        check(self.lib.cah_state_counts(st.handle, other.handle if other is not None else None,
                                        counts.ctypes.data, eq.ctypes.data if other is not None else None))
        return counts, (int(eq[0]) if other is not None else None)

    # state_counts(current_state(backend, st)) (src/measures.jl:14-20), counted on
    # the device (Int8 states)
    def state_counts(self, st):
        return count_dict(self._counts(st)[0])

    # cohens_kappa(current_state(st), current_state(observed)) (src/evaluation.jl:30-50)
    # from device counts; `observed` is a state of the same lattice
    def cohens_kappa(self, st, observed):
        cp, eq = self._counts(st, observed)
        co, _ = self._counts(observed)
        return MS.kappa_from_counts(count_dict(cp), count_dict(co), eq, int(np.prod(st.dims)))

    def state_like(self, prob, scheme, u):
        """A state of prob's lattice and rule holding `u` (an observed map to
        compare trajectories with)."""
        d, _keep = make_desc(prob, scheme, self.lenia_field)
        h = ctypes.c_void_p()
        check(self.lib.cah_state_create(self.ctx(), ctypes.byref(d), ctypes.byref(h)))
        st = HIPState(h, prob.space.dims, prob.u0.dtype)
        self.set_state(st, u)
        return st

    # -------------------------------------------------------------- batches

    def solve_batch(self, probs, scheme, use_graph=True, output="state", observed=None, saveat=None,
                    autotune=False, tile=None):
        """Several problems of one Int8 rule, advanced together.  output="state":
        their final states; "counts": state_counts of each final state; "kappa":
        cohens_kappa of each final state against `observed` (a map of the
        problems' common dims).  The last two are counted on the device, so only
        counters leave it.  With `saveat` (a step or a collection of times, as in
        solve) each problem's result is the list of that output at every saved
        time instead, t0 first: the batch advances between the saved times and
        the summaries are counted there on the device, without copying a lattice
        back.  With `autotune` the batch first times its band walk's candidate
        rows per wave over a few steps (cah_batch_autotune, which undoes them) and
        keeps the fastest; `self.batch_tile` is the choice, (dwords a lane, rows a
        wave), or (0, 0) when there is nothing to tune.  `tile` = (dwords a lane, rows
        a wave) sets the tile instead (cah_batch_tune)."""
        from .problem import _save_times
        if output not in ("state", "counts", "kappa"):
            raise M.ArgumentError(f"unknown output {output!r}")
        if output == "kappa" and observed is None:
            raise M.ArgumentError("output='kappa' needs the observed map")
        descs, keeps = [], []
        t0, t1 = probs[0].tspan
        for pr in probs:
            if pr.tspan != probs[0].tspan:
                raise M.ArgumentError("a batch advances problems over one tspan")
            d, k = make_desc(pr, scheme, self.lenia_field)
            descs.append(d)
            keeps.append(k)
        times = [t0, t1] if saveat is None else _save_times(t0, t1, saveat, False)
        arr = (L.CahDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        check(self.lib.cah_batch_create(self.ctx(), len(descs), arr, ctypes.byref(h)))
        obs = None
        try:
            for i, pr in enumerate(probs):
                u = np.asfortranarray(pr.u0)
                check(self.lib.cah_batch_set(h, i, u.ctypes.data, u.nbytes))
            if output == "kappa":
                obs = self.state_like(probs[0], scheme, observed)
            cp = make_params(probs[0], probs[0].p)
            if tile is not None:
                check(self.lib.cah_batch_tune(h, int(tile[0]), int(tile[1])))
            if autotune and t1 > t0:
                vb, rows = ctypes.c_int32(), ctypes.c_int32()
                check(self.lib.cah_batch_autotune(h, t0, min(20, t1 - t0), ctypes.byref(cp), ctypes.byref(vb),
                                                  ctypes.byref(rows)))
                self.batch_tile = (vb.value, rows.value)
            series, t = [], times[0]
            for k, ts in enumerate(times):
                if ts > t:
                    check(self.lib.cah_batch_advance(h, t, ts - t, ctypes.byref(cp), int(use_graph)))
                    t = ts
                if saveat is None and k == 0:
                    continue  # only the final result
                if output == "state":
                    outs = []
                    for i, pr in enumerate(probs):
                        out = np.empty(pr.space.dims, dtype=pr.u0.dtype, order="F")
                        check(self.lib.cah_batch_get(h, i, out.ctypes.data, out.nbytes))
                        outs.append(out)
                    series.append(outs)
                else:
                    series.append(self._batch_summaries(h, probs, output, obs))
            if saveat is None:
                return series[-1]
            return [[s[i] for s in series] for i in range(len(probs))]
        finally:
            if obs is not None:
                self.free_state(obs)
            self.lib.cah_batch_destroy(h)

    def _batch_summaries(self, h, probs, output, obs):
        """counts or kappa of every member of batch h; obs: the observed map's state."""
        n = len(probs)
        counts = np.zeros((n, 256), np.int64)
        eq = np.zeros(n, np.int64)
        check(self.lib.cah_batch_counts(h, obs.handle if obs is not None else None, counts.ctypes.data,
                                        eq.ctypes.data if obs is not None else None))
        if output == "counts":
            return [count_dict(c) for c in counts]
        co = count_dict(self._counts(obs)[0])
        cells = int(np.prod(obs.dims))
        return [MS.kappa_from_counts(count_dict(c), co, int(e), cells) for c, e in zip(counts, eq)]

    # -------------------------------------------------------------- launch checks

    def check_launches(self, prob, scheme=None, t0=None, nsteps=None):
        """The host-side launch checks (cah_check_launches) of prob's launches:
        every kernel launch of `nsteps` steps from t0 (the whole tspan by
        default), checked against the buffers it may read and write, without a
        device.  Returns {kernel: (launches, checks)}; raises DeviceError naming
        the first failed check."""
        scheme = M.Synchronous() if scheme is None else scheme
        d, _keep = make_desc(prob, scheme, self.lenia_field)
        t0 = prob.tspan[0] if t0 is None else t0
        nsteps = prob.tspan[1] - prob.tspan[0] if nsteps is None else nsteps
        cp = make_params(prob, prob.p)
        rep = ctypes.create_string_buffer(1 << 14)
        check(self.lib.cah_check_launches(ctypes.byref(d), int(t0), int(nsteps), ctypes.byref(cp), rep, len(rep)))
        return _parse_report(rep.value.decode())

    def check_batch_launches(self, probs, scheme=None, tile=(0, 0)):
        """The launch checks (cah_check_batch_launches) of a batch of probs over
        their common tspan, with the band walk's tile (dwords a lane, rows a wave;
        0 for the default)."""
        scheme = M.Synchronous() if scheme is None else scheme
        descs, keeps = zip(*[make_desc(pr, scheme, self.lenia_field) for pr in probs])
        arr = (L.CahDesc * len(descs))(*descs)
        t0, t1 = probs[0].tspan
        cp = make_params(probs[0], probs[0].p)
        rep = ctypes.create_string_buffer(1 << 14)
        check(self.lib.cah_check_batch_launches(len(descs), arr, int(t0), int(t1 - t0), ctypes.byref(cp),
                                                int(tile[0]), int(tile[1]), rep, len(rep)))
        return _parse_report(rep.value.decode())

    def __del__(self):
        if getattr(self, "_ctx", None) is not None:
            try:
                self.lib.cah_close(self._ctx)
            except Exception:
                pass


class HIPSparseBackend(HIPBackend):
    """HIPSparseBackend(background; tilesize=32): SparseBackend(background;
    tilesize) (src/sparse.jl:6-40) on the device.  Each step updates only the
    tiles within reach of a tile holding a non-background cell; the states and
    results are those of the dense sweep.  Synchronous and Stochastic schemes, no
    kernel fields, quiescent rules only (checked when the state is allocated, and
    at every step for the forest fire's parameters)."""

    def __init__(self, background, tilesize=32, device=0):
        super().__init__(device)
        if tilesize < 1:
            raise M.ArgumentError(f"tilesize must be at least 1, got {tilesize}")
        self.background = int(background)
        self.tilesize = int(tilesize)

    def supports(self, scheme):
        return isinstance(scheme, (M.Synchronous, M.Stochastic))

    def check_problem(self, prob, scheme):
        if isinstance(prob.rule, (M.Lenia, M.FieldEpidemic)):
            raise M.ArgumentError("SparseBackend does not support kernel fields; use ArrayBackend")
        super().check_problem(prob, scheme)

    def allocate_state(self, prob, scheme):
        st = super().allocate_state(prob, scheme)
        try:
            check(self.lib.cah_state_sparse(st.handle, self.background, self.tilesize))
        except Exception:
            self.free_state(st)
            raise
        return st

    # stored_tiles(integ) (src/sparse.jl)
    def stored_tiles(self, st):
        n = ctypes.c_int64()
        check(self.lib.cah_state_active_tiles(st.handle, ctypes.byref(n)))
        return n.value
