"""Summaries of states that ensembles and calibration reduce trajectories to.

    state_counts(u)                   src/measures.jl:14-20
    cohens_kappa(predicted, observed) src/evaluation.jl:30-50

Both take host arrays here.  Their device forms, which read HIP states and
bring back only counters, are HIPBackend.state_counts / .cohens_kappa and
solve_batch(..., output="counts" | "kappa"); they use kappa_from_counts, so a
device kappa is the same float the host form gives for the same maps.
"""
import numpy as np

from . import model as M


def state_counts(u):
    """How many cells hold each value of `u`: {value: count} over the values
    present (src/measures.jl:14-20)."""
    vals, cnt = np.unique(np.asarray(u), return_counts=True)
    return {v.item(): int(c) for v, c in zip(vals, cnt)}


def kappa_from_counts(cp, co, equal, n):
    """Cohen's kappa from the category counts of both maps, the number of cells
    where they agree and the cell count, with the reference's arithmetic
    (src/evaluation.jl:43-50): p_o = equal / n; p_e = (sum over observed
    categories of cp[k] * co[k], an integer) / n^2, each operand converted to
    Float64 before the division."""
    if n <= 0:
        raise M.ArgumentError("cohens_kappa: empty maps")
    po = float(equal) / float(n)
    pe = float(sum(cp.get(k, 0) * v for k, v in co.items())) / float(n * n)
    if pe == 1:
        return 1.0 if po == 1 else 0.0
    return (po - pe) / (1 - pe)


def cohens_kappa(predicted, observed):
    """Cohen's kappa between two categorical maps of the same size
    (src/evaluation.jl:30-50)."""
    p, o = np.asarray(predicted), np.asarray(observed)
    if p.shape != o.shape:
        raise M.ArgumentError(f"cohens_kappa: arrays have sizes {p.shape} and {o.shape}")
    return kappa_from_counts(state_counts(p), state_counts(o), int(np.count_nonzero(p == o)), o.size)
