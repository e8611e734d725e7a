"""cahip: the MI355X (gfx950) backend for the synchronous stencil sweep of
CellularAutomataCore.jl, driven through the C ABI of libcahip.so.

The host mirror of the reference's interface: model types (cahip.model), the
problem / solve callers (cahip.problem) and the backend (cahip.backend).
"""
from .measures import cohens_kappa, state_counts
from .model import (AnisotropicKernel, ArgumentError, CellularAutomataCoreError, Checkerboard,
                    Constant, DeviceError, EmpiricalKernel, ExponentialKernel, FieldEpidemic,
                    ForestFire, GaussianKernel, Generations, Grid, IncompatibleSpaceError,
                    InvalidStateError, Lattice, Lenia, LeniaKernel, Life, Margolus, Moore,
                    Offsets, OutOfMemoryError, Periodic, PowerLawKernel, RandomSequential,
                    Reflecting, RuleParseError, StateShapeError, Stochastic, Synchronous, Torus,
                    UniformKernel, UnsupportedCombinationError, VonNeumann, kernel_weights,
                    map_axis, neighbourhood_offsets)
from .problem import (CAProblem, CASolution, EnsembleProblem, final_state, seed_trajectory,
                      solve, solve_ensemble)


def HIPBackend(*args, **kw):
    """The HIP backend (loads libcahip.so; raises if it is not built)."""
    from .backend import HIPBackend as _B
    return _B(*args, **kw)
