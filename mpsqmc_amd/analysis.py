"""Statistics of a walk's energy trace.

The reference writes one row per step (``step N E_P E_T Omega``,
MPSQMC2.cpp:749-758) and leaves the analysis to the user.  The projected
energies of consecutive steps are correlated, so the error of their mean comes
from a blocking transformation (Flyvbjerg and Petersen, J. Chem. Phys. 91, 461
(1989)): average neighbouring pairs repeatedly; the variance of the mean
estimated at each level grows until the blocks are longer than the
correlation time and then stays on a plateau.
"""
import numpy as np


def blocking_levels(x):
    """Per blocking level: (block length, number of blocks, variance of the
    mean, its standard error).  Level 0 treats the samples as independent."""
    x = np.asarray(x, dtype=np.float64)
    out = []
    length = 1
    while x.size >= 2:
        n = x.size
        v = x.var(ddof=1) / n
        out.append((length, n, v, v * np.sqrt(2.0 / (n - 1))))
        if n < 4:
            break
        m = n // 2
        x = 0.5 * (x[0:2 * m:2] + x[1:2 * m:2])
        length *= 2
    return out


def blocking(x, min_blocks=16):
    """Mean, standard error and variance of the mean of a correlated series.

    The plateau is the first level that no later level (with at least
    ``min_blocks`` blocks) exceeds by more than the two standard errors
    combined."""
    x = np.asarray(x, dtype=np.float64)
    if x.size < 2:
        raise ValueError("blocking needs at least two samples")
    levels = [l for l in blocking_levels(x) if l[1] >= min_blocks] or blocking_levels(x)[:1]
    pick = levels[-1]
    for i, (length, n, v, dv) in enumerate(levels):
        if all(v2 - v <= dv + dv2 for (_, _, v2, dv2) in levels[i + 1:]):
            pick = levels[i]
            break
    length, n, v, dv = pick
    return {"mean": float(x.mean()), "error": float(np.sqrt(v)), "var_mean": float(v),
            "var_rel_error": float(dv / v) if v > 0 else float("inf"),
            "block_length": int(length), "n": int(x.size)}


def variance_time(energies, seconds, min_blocks=16):
    """Variance of the mean energy times the compute time that produced it:
    the cost of a run at fixed statistical error is proportional to it, so
    two proposals are compared by this product at equal energy error."""
    b = blocking(energies, min_blocks)
    b["seconds"] = float(seconds)
    b["variance_x_seconds"] = b["var_mean"] * float(seconds)
    return b
