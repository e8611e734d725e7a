"""CAProblem / solve / ensembles: the callers of the backend interface.

A thin mirror of src/problem.jl, enough to drive the HIP backend the way the
reference's CAIntegrator does: `solve` asks the backend `supports`, runs
`check_problem`, `allocate_state`, then `advance` between save points and
`current_state` at each of them (src/problem.jl:395-462, :557-559).
States are numpy arrays of shape `dims` in Fortran (column-major) order, so
u[i1, i2] is Julia's u[i1+1, i2+1].
"""
import numpy as np

from . import model as M


def _tspan(ts):
    if isinstance(ts, (int, np.integer)):
        return (0, int(ts))
    if isinstance(ts, tuple) and len(ts) == 2 and all(isinstance(x, (int, np.integer)) for x in ts):
        return (int(ts[0]), int(ts[1]))
    raise M.ArgumentError(f"tspan must be a number of steps or a tuple (t0, t1), got {ts}")


class CAProblem:
    """CAProblem(rule, u0, tspan, p=None; space, seed, mask, covariates)
    (src/problem.jl:67-92).  covariates: {name: array} of arrays the size of the
    space, or with one more axis of time slices (src/environment.jl:78-113)."""

    def __init__(self, rule, u0, tspan, p=None, space=None, seed=0, mask=None, covariates=None):
        u0 = np.asarray(u0)
        if u0.ndim == 0:
            raise M.ArgumentError("u0 must be an array")
        space = M.Torus(*u0.shape) if space is None else space
        ts = _tspan(tspan)
        if ts[1] < ts[0]:
            raise M.ArgumentError(f"tspan must not run backwards, got {ts}")
        if tuple(u0.shape) != space.dims:
            raise M.StateShapeError(f"the initial state: expected size {space.dims}, got {u0.shape}")
        if space.ndims not in (1, 2):
            raise M.IncompatibleSpaceError(f"{space}: the HIP backend runs 1-D and 2-D lattices")
        self.rule, self.space, self.tspan, self.p = rule, space, ts, p
        self.seed = int(seed) % (1 << 64)
        nb = rule.neighbourhood if rule.neighbourhood is not None else M.default_neighbourhood(space)
        self.neighbours = M.neighbourhood_offsets(nb, space)
        if isinstance(rule, M.Lenia):
            u0 = np.asarray(u0, dtype=np.float32)
        else:
            u0 = np.asarray(u0, dtype=np.int8)
            for b in space.boundaries:
                if isinstance(b, M.Constant):
                    v = b.value
                    if not (float(v) == int(v) and -128 <= int(v) <= 127):
                        raise M.IncompatibleSpaceError(
                            f"the boundary Constant({v!r}) on {space}: a state of Int8 cannot hold it")
        rule.validate_state(u0)
        self.u0 = np.asfortranarray(u0.copy())
        if mask is not None:
            mask = np.asarray(mask)
            if mask.dtype != np.bool_:
                raise M.ArgumentError("mask must be an array of Bools")
            if tuple(mask.shape) != space.dims:
                raise M.StateShapeError(f"the mask: expected size {space.dims}, got {mask.shape}")
            mask = np.asfortranarray(mask)
        self.mask = mask
        # the rule's own :potential KernelField (src/lenia.jl:36-39), resolved as
        # _resolve_field does: offsets to read are -o (src/environment.jl:59-69)
        self.taps = None
        if isinstance(rule, M.Lenia):
            offs, w = M.kernel_weights(rule.kernel, space)
            if rule.method == "fft" and not all(isinstance(b, M.Periodic) for b in space.boundaries):
                raise M.ArgumentError("field :potential uses method=:fft, which needs periodic boundaries on every axis")
            self.taps = ([(-a, -b) for a, b in offs], w)
        self.covariates = {}
        for name, a in (covariates or {}).items():
            a = np.asarray(a, dtype=np.float64)
            N = space.ndims
            if a.ndim not in (N, N + 1) or tuple(a.shape[:N]) != space.dims:
                raise M.StateShapeError(f"covariate :{name}: expected size {space.dims}, got {a.shape}")
            self.covariates[name] = np.asfortranarray(a)
        if isinstance(rule, M.FieldEpidemic):
            offs, w = M.kernel_weights(rule.kernel, space, include_centre=rule.include_centre)
            self.taps = ([tuple(-x for x in (o + (0,))[:2]) for o in offs], w)
            if rule.covariate is not None and rule.covariate not in self.covariates:
                raise M.ArgumentError(f"the rule reads covariate :{rule.covariate}, which the problem lacks")

    def remake(self, **kw):
        """remake(prob; ...) (src/problem.jl:103-113)."""
        args = dict(rule=self.rule, u0=self.u0, tspan=self.tspan, p=self.p, space=self.space,
                    seed=self.seed, mask=self.mask, covariates=self.covariates)
        args.update(kw)
        return CAProblem(**args)


class CASolution:
    """Saved times and states of a run (src/problem.jl:132-176)."""

    def __init__(self, t, u):
        self.t, self.u = t, u


def final_state(sol):
    return sol.u[-1]


def _save_times(t0, t1, saveat, save_everystep):
    """The saved times (src/problem.jl:335-363): t0, then every `saveat` steps or
    the times of a collection within (t0, t1], and t1."""
    if saveat is not None:
        if np.ndim(saveat) == 0:
            if int(saveat) <= 0:
                raise M.ArgumentError(f"saveat must be positive, got {saveat}")
            ts = list(range(t0, t1 + 1, int(saveat)))
        else:
            ts = [t0] + sorted({int(x) for x in saveat if t0 < x <= t1})
        if ts[-1] != t1:
            ts.append(t1)
        return ts
    if save_everystep:
        return list(range(t0, t1 + 1))
    return [t0, t1] if t1 > t0 else [t0]


def solve(prob, alg=None, backend=None, saveat=None, save_everystep=None):
    """solve(prob, alg; backend, saveat, save_everystep) (src/problem.jl:395-499, :557-559)."""
    from .backend import HIPBackend
    alg = M.Synchronous() if alg is None else alg
    backend = HIPBackend() if backend is None else backend
    if not backend.supports(alg):
        raise M.UnsupportedCombinationError(f"{type(backend).__name__} does not support {type(alg).__name__}")
    backend.check_problem(prob, alg)
    st = backend.allocate_state(prob, alg)
    try:
        t0, t1 = prob.tspan
        everystep = saveat is None if save_everystep is None else save_everystep
        times = _save_times(t0, t1, saveat, everystep)
        us, t = [backend.current_state(st)], t0
        for ts in times[1:]:
            backend.advance(st, prob, alg, prob.p, t, ts - t)
            t = ts
            us.append(backend.current_state(st))
        return CASolution(times, us)
    finally:
        backend.free_state(st)


def seed_trajectory(prob, i):
    """seed_trajectory (src/problem.jl:571-574): trajectory i (1-based) gets seed prob.seed + i."""
    return prob.remake(seed=(prob.seed + i) % (1 << 64))


class EnsembleProblem:
    """EnsembleProblem(prob; prob_func=seed_trajectory) (src/problem.jl:585-589)."""

    def __init__(self, prob, prob_func=seed_trajectory):
        self.prob, self.prob_func = prob, prob_func


def solve_ensemble(eprob, alg=None, backend=None, trajectories=1, batched=True, saveat=None):
    """solve(eprob, alg, EnsembleThreads(); trajectories) for the final states.

    With `batched` the trajectories advance together through the backend's
    batched form (one launch per step for all of them); otherwise each runs its
    own plain sweep, as independent integrators would.  Returns the list of
    final states, trajectory 1 first; with `saveat`, the list of each
    trajectory's states at the saved times (as solve's `u`).
    """
    from .backend import HIPBackend
    alg = M.Synchronous() if alg is None else alg
    backend = HIPBackend() if backend is None else backend
    probs = [eprob.prob_func(eprob.prob, i) for i in range(1, trajectories + 1)]
    if batched:
        return backend.solve_batch(probs, alg, saveat=saveat)
    if saveat is not None:
        return [solve(p, alg, backend=backend, saveat=saveat).u for p in probs]
    return [final_state(solve(p, alg, backend=backend, save_everystep=False)) for p in probs]
