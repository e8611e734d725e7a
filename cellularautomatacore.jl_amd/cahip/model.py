"""The model vocabulary the HIP backend runs, mirroring CellularAutomataCore.jl.

Spaces and boundaries (src/spaces.jl), neighbourhoods (src/neighbourhoods.jl),
the rule families on the sweep (src/rules.jl, src/lenia.jl, src/kernels.jl),
update schemes (src/schemes.jl) and the exception hierarchy (src/errors.jl).
Names, argument meanings and error types follow the reference so that tests
read like its own.
"""
import math
import re

import numpy as np

# ------------------------------------------------------------------ errors (src/errors.jl)


class CellularAutomataCoreError(Exception):
    """Root of the package's exceptions (src/errors.jl:28)."""


class ArgumentError(CellularAutomataCoreError, ValueError):
    """Julia's ArgumentError, raised for invalid arguments."""


class RuleParseError(CellularAutomataCoreError, ValueError):
    """A rule string that cannot be parsed (src/errors.jl)."""


class StateShapeError(CellularAutomataCoreError, ValueError):
    """A state of the wrong size for its space (src/errors.jl)."""


class InvalidStateError(CellularAutomataCoreError, ValueError):
    """A state value the rule cannot read (src/errors.jl)."""


class IncompatibleSpaceError(CellularAutomataCoreError, ValueError):
    """A component that does not fit the space (src/errors.jl:108-116)."""


class UnsupportedCombinationError(CellularAutomataCoreError):
    """A backend that cannot run a scheme (src/errors.jl:124-132)."""


class DeviceError(CellularAutomataCoreError, RuntimeError):
    """A HIP runtime failure."""


class OutOfMemoryError(CellularAutomataCoreError, MemoryError):
    """Julia's OutOfMemoryError: device memory exhausted."""


# ------------------------------------------------------------------ spaces (src/spaces.jl)


class Periodic:
    """The axis wraps around (src/spaces.jl:16)."""
    code = 0

    def __eq__(self, o):
        return isinstance(o, Periodic)

    def __repr__(self):
        return "Periodic()"


class Reflecting:
    """Mirrored at the edges with the edge cell repeated (src/spaces.jl:25)."""
    code = 1

    def __eq__(self, o):
        return isinstance(o, Reflecting)

    def __repr__(self):
        return "Reflecting()"


class Constant:
    """Every cell beyond the edge holds `value` (src/spaces.jl:40-42)."""
    code = 2

    def __init__(self, value=0):
        self.value = value

    def __eq__(self, o):
        return isinstance(o, Constant) and o.value == self.value

    def __repr__(self):
        return f"Constant({self.value!r})"


def map_axis(b, j, n):
    """_map_axis (src/spaces.jl:46-51): 1-based j -> (index, outside)."""
    if isinstance(b, Periodic):
        return ((j - 1) % n) + 1, False
    if isinstance(b, Reflecting):
        m = (j - 1) % (2 * n)
        return (m + 1 if m < n else 2 * n - m), False
    return j, not (1 <= j <= n)


class Lattice:
    """A square lattice with one boundary per axis (src/spaces.jl:94-110)."""

    def __init__(self, dims, boundaries):
        dims = tuple(int(d) for d in dims)
        if not all(d > 0 for d in dims):
            raise ArgumentError(f"lattice dimensions must be positive, got {dims}")
        if len(boundaries) != len(dims):
            raise ArgumentError(f"need one boundary per axis ({len(dims)}), got {len(boundaries)}")
        self.dims = dims
        self.boundaries = tuple(boundaries)

    @property
    def ndims(self):
        return len(self.dims)

    def __len__(self):
        return int(np.prod(self.dims))

    def __repr__(self):
        return f"Lattice({self.dims}, {self.boundaries})"


def Torus(*dims):
    """A lattice periodic on every axis (src/spaces.jl:130-131)."""
    if len(dims) == 1 and isinstance(dims[0], tuple):
        dims = dims[0]
    return Lattice(dims, tuple(Periodic() for _ in dims))


def Grid(*dims, boundary=None):
    """A lattice with edges, Constant(0) by default (src/spaces.jl:139-143)."""
    if len(dims) == 1 and isinstance(dims[0], tuple):
        dims = dims[0]
    boundary = Constant(0) if boundary is None else boundary
    bs = tuple(boundary) if isinstance(boundary, tuple) else tuple(boundary for _ in dims)
    return Lattice(dims, bs)


# ------------------------------------------------------------------ neighbourhoods


class Moore:
    """Cells within Chebyshev distance `radius` (src/neighbourhoods.jl:21-27)."""

    def __init__(self, radius=1):
        if radius < 1:
            raise ArgumentError(f"Moore radius must be at least 1, got {radius}")
        self.radius = int(radius)

    def __eq__(self, o):
        return isinstance(o, Moore) and o.radius == self.radius

    def __repr__(self):
        return f"Moore({self.radius})"


class VonNeumann:
    """Cells within Manhattan distance `radius` (src/neighbourhoods.jl:35-42)."""

    def __init__(self, radius=1):
        if radius < 1:
            raise ArgumentError(f"von Neumann radius must be at least 1, got {radius}")
        self.radius = int(radius)

    def __eq__(self, o):
        return isinstance(o, VonNeumann) and o.radius == self.radius

    def __repr__(self):
        return f"VonNeumann({self.radius})"


class Offsets:
    """Explicit offsets in a given order (src/neighbourhoods.jl:79-97)."""

    def __init__(self, *offsets):
        offs = tuple(tuple(int(x) for x in o) for o in offsets)
        if not offs:
            raise ArgumentError("a neighbourhood needs at least one offset")
        if any(all(x == 0 for x in o) for o in offs):
            raise ArgumentError("the zero offset (the centre cell) cannot be a neighbour")
        if len(set(offs)) != len(offs):
            raise ArgumentError("offsets must be unique")
        self.offsets = offs

    def __len__(self):
        return len(self.offsets)

    def __iter__(self):
        return iter(self.offsets)

    def __eq__(self, o):
        return isinstance(o, Offsets) and o.offsets == self.offsets

    def __repr__(self):
        return "Offsets" + repr(self.offsets)


def _offsets_where(pred, N, r):
    # CartesianIndices((-r:r)^N), first axis fastest (src/neighbourhoods.jl:124-128)
    out = []
    if N == 1:
        for o1 in range(-r, r + 1):
            if o1 != 0 and pred((o1,)):
                out.append((o1,))
    else:
        for o2 in range(-r, r + 1):
            for o1 in range(-r, r + 1):
                if (o1, o2) != (0, 0) and pred((o1, o2)):
                    out.append((o1, o2))
    return Offsets(*out)


def neighbourhood_offsets(n, space):
    """neighbourhood_offsets (src/neighbourhoods.jl:148-174)."""
    N = space.ndims
    if isinstance(n, Moore):
        return _offsets_where(lambda o: True, N, n.radius)
    if isinstance(n, VonNeumann):
        return _offsets_where(lambda o: sum(abs(x) for x in o) <= n.radius, N, n.radius)
    if isinstance(n, Offsets):
        if any(len(o) != N for o in n.offsets):
            raise IncompatibleSpaceError(
                f"{n} on {space}: offsets do not have the lattice's {N} dimensions")
        return n
    raise IncompatibleSpaceError(f"{n} does not fit {space}")


def default_neighbourhood(space):
    """Moore(1) on a square lattice (src/neighbourhoods.jl:120)."""
    return Moore(1)


def halo(offsets):
    """_halo (src/neighbourhoods.jl:177): the largest absolute offset."""
    return max(max(abs(x) for x in o) for o in offsets)


# ------------------------------------------------------------------ rules (src/rules.jl)

LIFE_PRESETS = {"conway": "B3/S23", "highlife": "B36/S23", "seeds": "B2/S",
                "day_and_night": "B3678/S34678", "morley": "B368/S245",
                "two_by_two": "B36/S125", "replicator": "B1357/S1357",
                "diamoeba": "B35678/S5678", "thirty_four": "B34/S34",
                "life_without_death": "B3/S012345678", "maze": "B3/S12345"}
GENERATIONS_PRESETS = {"brians_brain": "B2/S/C3", "star_wars": "B2/S345/C4"}


def _mask(counts, rule):
    m = 0
    for c in counts:
        c = int(c)
        if not 0 <= c <= 63:
            raise ArgumentError(f"neighbour counts must be in 0:63, got {c} in {rule}")
        m |= 1 << c
    return m


def _parse_digits(s, rule):
    for ch in s:
        if not "0" <= ch <= "8":
            raise RuleParseError(f"{rule}: neighbour counts in a rule string are digits 0 to 8, got {ch!r}")
    return [ord(ch) - 48 for ch in s]


def _parse_rule(text):
    """_parse_rule (src/rules.jl:316-330): (birth digits, survive digits, C or None)."""
    s = re.sub(r"\s", "", text).upper()
    m = re.match(r"^B([0-9]*)/?S([0-9]*)(?:/C?([0-9]+))?$", s)
    if m:
        c = None if m.group(3) is None else int(m.group(3))
        return _parse_digits(m.group(1), text), _parse_digits(m.group(2), text), c
    m = re.match(r"^([0-9]*)/([0-9]*)(?:/([0-9]+))?$", s)
    if m:
        c = None if m.group(3) is None else int(m.group(3))
        return _parse_digits(m.group(2), text), _parse_digits(m.group(1), text), c
    raise RuleParseError(f"{text}: expected B/S notation such as \"B3/S23\" or S/B notation such as \"23/3\"")


def _digitstring(m):
    cs = [c for c in range(64) if (m >> c) & 1]
    if all(c <= 8 for c in cs):
        return "".join(str(c) for c in cs)
    runs, i = [], 0
    while i < len(cs):
        j = i
        while j + 1 < len(cs) and cs[j + 1] == cs[j] + 1:
            j += 1
        runs.append(f"{cs[i]}..{cs[j]}" if j > i else str(cs[i]))
        i = j + 1
    return ",".join(runs)


class Life:
    """A two-state outer-totalistic rule (src/rules.jl:354-389).

    Life("B3/S23"), Life("conway") (a preset), or Life(birth_counts, survive_counts).
    """
    code = 1

    def __init__(self, rule="B3/S23", survive=None, neighbourhood=None):
        if survive is not None:
            self.birth, self.survive = _mask(rule, "Life"), _mask(survive, "Life")
        else:
            # a preset name stands for Julia's Symbol form, Life(:conway)
            rule = LIFE_PRESETS.get(rule, rule)
            b, s, c = _parse_rule(rule)
            if c is not None and c != 2:
                raise RuleParseError(f"{rule}: a Life rule has two states; use Generations for C{c}")
            self.birth, self.survive = _mask(b, rule), _mask(s, rule)
        self.neighbourhood = neighbourhood

    def rulestring(self):
        return "B" + _digitstring(self.birth) + "/S" + _digitstring(self.survive)

    def validate_state(self, u0):
        if not np.isin(u0, (0, 1)).all():
            raise InvalidStateError("a Life rule has states 0 and 1")

    def __repr__(self):
        return f"Life({self.rulestring()!r})"


class Generations:
    """A Generations rule with C states (src/rules.jl:424-456)."""
    code = 2

    def __init__(self, rule, neighbourhood=None):
        if rule in GENERATIONS_PRESETS:
            rule = GENERATIONS_PRESETS[rule]
        b, s, c = _parse_rule(rule)
        if c is None:
            raise RuleParseError(f"{rule}: a Generations rule needs a state count, as in \"B2/S/C3\"")
        if c < 2:
            raise RuleParseError(f"{rule}: a Generations rule needs at least 2 states")
        self.birth, self.survive, self.states = _mask(b, rule), _mask(s, rule), c
        self.neighbourhood = neighbourhood

    def rulestring(self):
        return ("B" + _digitstring(self.birth) + "/S" + _digitstring(self.survive)
                + "/C" + str(self.states))

    def validate_state(self, u0):
        if not ((u0 >= 0) & (u0 < self.states)).all():
            raise InvalidStateError(f"states must be in 0:{self.states - 1}")


class ForestFire:
    """The Drossel-Schwabl forest fire of README.md:56-60 as an Int8 rule.

    The reference writes it as a LocalRule closure over Symbol states
    (:empty, :tree, :fire); here the states are Int8 0, 1, 2 under that
    bijection, and the rule body is the same: a burning tree burns out, empty
    ground grows a tree with probability p.p (draw k = 1), a tree next to a fire
    catches fire, any other tree is struck with probability p.f (draw k = 2).
    """
    code = 3
    EMPTY, TREE, FIRE = 0, 1, 2

    def __init__(self, neighbourhood=None):
        self.neighbourhood = VonNeumann(1) if neighbourhood is None else neighbourhood

    def validate_state(self, u0):
        if not np.isin(u0, (0, 1, 2)).all():
            raise InvalidStateError("forest-fire states are 0 (empty), 1 (tree) and 2 (fire)")

    def __repr__(self):
        return "ForestFire()"


# ------------------------------------------------------------------ dispersal kernels (src/kernels.jl)


def _positive_radius(radius):
    if radius < 1:
        raise ArgumentError(f"radius must be at least 1, got {radius}")
    return int(radius)


class GaussianKernel:
    """GaussianKernel(sigma; radius=ceil(3 sigma)): weight exp(-d^2 / 2 sigma^2)
    (src/kernels.jl:15-30)."""

    def __init__(self, sigma, radius=None):
        if not sigma > 0:
            raise ArgumentError(f"σ must be positive, got {sigma}")
        self.sigma = float(sigma)
        self.radius = _positive_radius(math.ceil(3 * self.sigma) if radius is None else radius)

    def value(self, d, o):
        # d^2 and σ^2 are x*x in Julia (literal powers)
        return math.exp(-(d * d) / (2 * (self.sigma * self.sigma)))


class ExponentialKernel:
    """ExponentialKernel(scale; radius=ceil(5 scale)): weight exp(-d / scale)
    (src/kernels.jl:32-47)."""

    def __init__(self, scale, radius=None):
        if not scale > 0:
            raise ArgumentError(f"scale must be positive, got {scale}")
        self.scale = float(scale)
        self.radius = _positive_radius(math.ceil(5 * self.scale) if radius is None else radius)

    def value(self, d, o):
        return math.exp(-d / self.scale)


class PowerLawKernel:
    """PowerLawKernel(alpha; radius=10): weight (1 + d)^(-alpha) (src/kernels.jl:49-64)."""

    def __init__(self, alpha, radius=10):
        if not alpha > 0:
            raise ArgumentError(f"α must be positive, got {alpha}")
        self.alpha = float(alpha)
        self.radius = _positive_radius(radius)

    def value(self, d, o):
        return (1 + d) ** (-self.alpha)


class UniformKernel:
    """UniformKernel(radius): equal weight within distance radius (src/kernels.jl:66-79)."""

    def __init__(self, radius):
        self.radius = _positive_radius(radius)

    def value(self, d, o):
        return 1.0


class LeniaKernel:
    """The smooth shell kernel of Lenia (src/kernels.jl:87-104)."""

    def __init__(self, R, peaks=(1.0,)):
        if R < 1:
            raise ArgumentError(f"R must be at least 1, got {R}")
        if len(peaks) == 0:
            raise ArgumentError("at least one peak is needed")
        self.radius = int(R)
        self.peaks = tuple(float(p) for p in peaks)

    def value(self, d, o=None):
        r = d / self.radius
        if r <= 0 or r >= 1:
            return 0.0
        B = len(self.peaks)
        x = B * r
        i = min(math.floor(x), B - 1)
        f = x - i
        return self.peaks[i] * math.exp(4 - 1 / (f * (1 - f))) if 0 < f < 1 else 0.0


class AnisotropicKernel:
    """AnisotropicKernel(base, direction, strength): the base weight times
    exp(strength * cos(theta - direction)), theta the angle of the offset with
    east = +column (second axis), north = -row (src/kernels.jl:106-131)."""

    def __init__(self, base, direction, strength):
        if not strength >= 0:
            raise ArgumentError(f"strength must be nonnegative, got {strength}")
        self.base = base
        self.direction = float(direction)
        self.strength = float(strength)
        self.radius = base.radius

    def value(self, d, o):
        w = self.base.value(d, o)
        if d == 0:
            return w
        theta = math.atan2(-o[0], o[1])
        return w * math.exp(self.strength * math.cos(theta - self.direction))


class EmpiricalKernel:
    """EmpiricalKernel(weights): a matrix of weights with odd sides centred on the
    cell (src/kernels.jl:133-150)."""

    def __init__(self, weights):
        w = np.array(weights, dtype=np.float64)
        if w.ndim != 2 or not all(n % 2 == 1 for n in w.shape):
            raise ArgumentError(f"kernel matrix sides must be odd, got {w.shape}")
        if not (np.isfinite(w).all() and (w >= 0).all()):
            raise ArgumentError("kernel weights must be finite and nonnegative")
        if not w.sum() > 0:
            raise ArgumentError("kernel weights must not all be zero")
        self.weights = w
        self.radius = max(n // 2 for n in w.shape)


def kernel_weights(kernel, space, include_centre=True):
    """kernel_weights (src/kernels.jl:174-212): the kernel's offsets, in
    column-major order, and their FP64 weights normalised to sum to one."""
    if not isinstance(space, Lattice):
        raise IncompatibleSpaceError(f"{type(kernel).__name__}: dispersal kernels need a lattice")
    offs, ws = [], []
    if isinstance(kernel, EmpiricalKernel):
        if space.ndims != 2:
            raise IncompatibleSpaceError("EmpiricalKernel: an empirical kernel matrix is two-dimensional")
        c1, c2 = (n // 2 for n in kernel.weights.shape)
        for j in range(kernel.weights.shape[1]):
            for i in range(kernel.weights.shape[0]):
                o = (i - c1, j - c2)
                if o == (0, 0) and not include_centre:
                    continue
                w = float(kernel.weights[i, j])
                if w > 0:
                    offs.append(o)
                    ws.append(w)
        if not ws:
            raise ArgumentError("the empirical kernel has no positive weight")
    else:
        r = kernel.radius
        if isinstance(kernel, AnisotropicKernel) and space.ndims != 2:
            raise IncompatibleSpaceError("AnisotropicKernel: directions need a two-dimensional lattice")
        grid = ([(o1,) for o1 in range(-r, r + 1)] if space.ndims == 1 else
                [(o1, o2) for o2 in range(-r, r + 1) for o1 in range(-r, r + 1)])
        for o in grid:
            if all(x == 0 for x in o) and not include_centre:
                continue
            d = math.sqrt(sum(x * x for x in o))
            if not d <= r + 1e-9:
                continue
            w = float(kernel.value(d, o))
            if w > 0:
                offs.append(o)
                ws.append(w)
        if not ws:
            raise ArgumentError(f"kernel {type(kernel).__name__} has no positive weight on {space}")
    w = np.asarray(ws, dtype=np.float64)
    return offs, w / w.sum()


# ------------------------------------------------------------------ the field epidemic


class FieldEpidemic:
    """FieldEpidemic(kernel; include_centre=True, precision=64, form="bernoulli",
    recover=False, covariate=None): an SI (or SIR) epidemic on Int8 states S=0,
    I=1, R=2 driven by a KernelField whose source is the infected cells,
    KernelField(kernel, ==(1); include_centre, precision) (src/environment.jl:5-36).

    A susceptible cell is infected with probability q = 1 - exp(-beta * c * field),
    with c the covariate named `covariate` (1 without one) and beta, gamma the
    problem's parameters p.beta, p.gamma.  form="bernoulli" is the LocalRule
    `bernoulli(ctx, q) ? 1 : 0` (test/backend_conformance.jl:63-68, :82-85);
    form="stated" is the ProbabilisticRule with outcomes (0 => 1 - q, 1 => q)
    (:101-104).  With recover=True an infected cell is removed with probability
    gamma (draw k = 2, or the stated outcomes (1 => 1 - gamma, 2 => gamma)).  The
    arithmetic is Float32 when beta (numpy.float32), the field (precision=32) are
    and there is no covariate, else Float64."""
    code = 5

    def __init__(self, kernel, include_centre=True, precision=64, form="bernoulli",
                 recover=False, covariate=None):
        if precision not in (32, 64):
            raise ArgumentError("precision must be 32 or 64")
        if form not in ("bernoulli", "stated"):
            raise ArgumentError(f"form must be 'bernoulli' or 'stated', got {form!r}")
        self.kernel = kernel
        self.include_centre = bool(include_centre)
        self.precision = precision
        self.form = form
        self.recover = bool(recover)
        self.covariate = covariate
        self.neighbourhood = None

    def validate_state(self, u0):
        allowed = (0, 1, 2) if self.recover else (0, 1)
        if not np.isin(u0, allowed).all():
            raise InvalidStateError(f"allowed states are {list(allowed)}")

    def __repr__(self):
        return (f"FieldEpidemic({type(self.kernel).__name__}, form={self.form!r}, "
                f"recover={self.recover})")


# ------------------------------------------------------------------ Lenia (src/lenia.jl)


class Lenia:
    """Lenia(; R=13, mu=0.15, sigma=0.015, dt=0.1, peaks=(1.0,), method="direct",
    precision=64) (src/lenia.jl:17-32)."""
    code = 4

    def __init__(self, R=13, mu=0.15, sigma=0.015, dt=0.1, peaks=(1.0,), method="direct",
                 precision=64):
        if not sigma > 0:
            raise ArgumentError(f"sigma must be positive, got {sigma}")
        if not 0 < dt <= 1:
            raise ArgumentError(f"dt must be in (0, 1], got {dt}")
        if method not in ("direct", "fft"):
            raise ArgumentError(f"method must be :direct or :fft, got :{method}")
        if precision not in (32, 64):
            raise ArgumentError("precision must be 32 or 64")
        cast = (lambda v: float(np.float32(v))) if precision == 32 else float
        self.mu, self.sigma, self.dt = cast(mu), cast(sigma), cast(dt)
        self.kernel = LeniaKernel(R, peaks)
        self.method = method
        self.precision = precision
        self.neighbourhood = None

    def validate_state(self, u0):
        if not np.issubdtype(u0.dtype, np.floating):
            raise InvalidStateError("Lenia needs a floating-point state")
        if not ((u0 >= 0) & (u0 <= 1)).all():
            raise InvalidStateError("states must be in [0, 1]")

    def __repr__(self):
        return f"Lenia(R={self.kernel.radius}, mu={self.mu}, sigma={self.sigma}, dt={self.dt})"


# ------------------------------------------------------------------ schemes (src/schemes.jl)


class Synchronous:
    """All cells update at once (src/schemes.jl:22)."""
    code = 0


class Stochastic:
    """Each cell updates with probability p, decided by its draw k = 0 (src/schemes.jl:33-40).

    A numpy.float32 p uses the 24-bit draw, as Stochastic(0.5f0) does."""
    code = 1

    def __init__(self, p):
        if not 0 < float(p) <= 1:
            raise ArgumentError(f"update probability must be in (0, 1], got {p}")
        self.p = p
        self.is_f32 = isinstance(p, np.float32)


class Checkerboard:
    """Two half-steps: even index sums first, then odd (src/schemes.jl:51)."""
    code = 2


class RandomSequential:
    """Sequential scheme; not run by the HIP backend (src/schemes.jl:62)."""
    code = 100


class Margolus:
    """Block scheme; not run by the HIP backend (src/schemes.jl:72)."""
    code = 101
