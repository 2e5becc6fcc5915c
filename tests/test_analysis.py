"""Blocking analysis of correlated energy traces against closed forms."""
import numpy as np
import pytest
from scipy.signal import lfilter

from mpsqmc_amd.analysis import blocking, blocking_levels, variance_time


def _ar1(phi, n, rng):
    """x_t = phi x_{t-1} + e_t, started in its stationary distribution."""
    e = rng.normal(size=n)
    e[0] /= np.sqrt(1.0 - phi * phi)
    return lfilter([1.0], [1.0, -phi], e)


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9])
def test_blocking_matches_ar1_variance_of_mean(phi):
    # variance of the mean of a stationary AR(1) series of length n:
    # var(x)/n * (1 + phi)/(1 - phi), var(x) = 1/(1 - phi^2)
    rng = np.random.default_rng(11)
    n = 1 << 15
    exact = (1.0 / (1.0 - phi * phi)) / n * (1.0 + phi) / (1.0 - phi)
    ratios = [blocking(_ar1(phi, n, rng))["var_mean"] / exact for _ in range(12)]
    assert 0.8 < np.mean(ratios) < 1.2


def test_blocking_levels_halve_and_preserve_mean():
    x = np.arange(40, dtype=float)
    lv = blocking_levels(x)
    assert [l[1] for l in lv] == [40, 20, 10, 5, 2]
    assert [l[0] for l in lv] == [1, 2, 4, 8, 16]
    # level 0 is the naive estimate
    assert lv[0][2] == pytest.approx(x.var(ddof=1) / 40)


def test_variance_time_scales_with_seconds():
    rng = np.random.default_rng(5)
    x = rng.normal(size=4096)
    a = variance_time(x, 2.0)
    assert a["variance_x_seconds"] == pytest.approx(2.0 * a["var_mean"])
    assert a["mean"] == pytest.approx(x.mean())
    with pytest.raises(ValueError):
        blocking([1.0])
