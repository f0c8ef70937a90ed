"""Host-side analysis of measurement series: binned errors, integrated
autocorrelation time and independent configurations per second (the
"indep. configs/sec" figure of merit of ensemble generation)."""
import numpy as np


def binned_error(x, nbins=10):
    """Mean and standard error of the mean from `nbins` equal bins (excess
    samples at the end are dropped)."""
    x = np.asarray(x, dtype=float)
    n = len(x) // nbins * nbins
    if n < 2 * nbins:
        raise ValueError("need at least two samples per bin")
    b = x[:n].reshape(nbins, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbins)


def autocorrelation(x, tmax=None):
    """Normalised autocorrelation rho(t), t = 0 .. tmax, of a stationary
    series (FFT, zero-padded, biased estimator)."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    tmax = n - 1 if tmax is None else min(int(tmax), n - 1)
    d = x - x.mean()
    f = np.fft.rfft(d, 2 * n)
    c = np.fft.irfft(f * np.conj(f))[: tmax + 1] / n
    return c / c[0]


def tau_int(x, c=5.0):
    """Integrated autocorrelation time with automatic windowing (Madras and
    Sokal, J. Stat. Phys. 50 (1988) 109): tau(W) = 1/2 + sum_{t=1}^{W} rho(t)
    for the smallest W >= c tau(W).  Returns (tau, error, W) with the
    error estimate tau sqrt(2 (2W + 1) / N)."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    rho = autocorrelation(x)
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return tau, tau * np.sqrt(2.0 * (2 * w + 1) / n), w
    return tau, tau * np.sqrt(2.0 * (2 * (n - 1) + 1) / n), n - 1


def independent_configs_per_second(sweeps_per_second, series, c=5.0):
    """Sweeps per second divided by 2 tau_int of the series (one measurement
    per sweep)."""
    tau, _, _ = tau_int(series, c)
    return sweeps_per_second / (2.0 * tau)
