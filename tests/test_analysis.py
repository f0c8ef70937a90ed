"""tau_int and binned errors against closed forms: an AR(1) process
x_t = rho x_{t-1} + e_t has rho(t) = rho^t and
tau_int = (1 + rho) / (2 (1 - rho)); white noise has tau_int = 1/2."""
import numpy as np
import pytest

from qcdgpu_amd import analysis as A


def _ar1(rho, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


@pytest.mark.parametrize("rho", [0.5, 0.9, 0.97])
def test_tau_int_ar1(rho):
    x = _ar1(rho, 400_000, 1)
    tau, err, w = A.tau_int(x)
    exact = (1 + rho) / (2 * (1 - rho))
    assert abs(tau - exact) < max(4 * err, 0.02 * exact), (tau, exact, err, w)


def test_tau_int_white_noise():
    x = np.random.default_rng(2).normal(size=200_000)
    tau, err, _ = A.tau_int(x)
    assert abs(tau - 0.5) < 4 * err + 1e-3


def test_autocorrelation_ar1_shape():
    rho = 0.8
    r = A.autocorrelation(_ar1(rho, 400_000, 3), tmax=10)
    assert np.allclose(r, rho ** np.arange(11), atol=0.01)


def test_binned_error_matches_tau():
    """For a correlated series the binned error with long bins approaches
    sqrt(2 tau var / N)."""
    rho = 0.9
    x = _ar1(rho, 400_000, 4)
    _, err = A.binned_error(x, nbins=40)
    tau = (1 + rho) / (2 * (1 - rho))
    expected = np.sqrt(2 * tau * np.var(x) / len(x))
    assert abs(err / expected - 1) < 0.35


def test_independent_configs():
    x = _ar1(0.9, 200_000, 5)
    tau, _, _ = A.tau_int(x)
    assert A.independent_configs_per_second(1000.0, x) == pytest.approx(1000.0 / (2 * tau))
