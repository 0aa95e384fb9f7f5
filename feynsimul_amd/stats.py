"""Error analysis of per-chain measurement series (host side, FP64).

The reference reports a 95% interval mean +- 1.96 std / sqrt(walkers) over
walker means (examples/lm2m2_plotter.py:100-117).  Here the series of every
chain is kept, so the integrated autocorrelation time and the effective
sample size are available as well:

  tau_int = 1 + 2 sum_{t=1}^{W} rho(t),  W the first window with W >= c tau_int(W)
  (Sokal's automatic windowing, c = 5),  ESS = chains * measurements / tau_int.
"""
import numpy as np


def autocorrelation(series):
    """Normalised autocorrelation rho(t), t = 0..M-1, pooled over chains.

    series: (measurements, chains).  The chains sample one distribution, so
    they are centred on the grand mean (centring each chain on its own mean
    would bias tau low by about tau / measurements); the autocovariance is
    averaged over chains (FFT, zero padded).
    """
    x = np.asarray(series, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    M = x.shape[0]
    x = x - x.mean()
    n = 1 << (2 * M - 1).bit_length()
    f = np.fft.rfft(x, n=n, axis=0)
    acov = np.fft.irfft(f * np.conj(f), n=n, axis=0)[:M].mean(axis=1) / M
    if acov[0] <= 0:
        return np.ones(1)
    return acov / acov[0]


def integrated_autocorr_time(series, c=5.0):
    """(tau_int, window) with Sokal's windowing."""
    rho = autocorrelation(series)
    tau = 1.0
    for W in range(1, len(rho)):
        tau = 1.0 + 2.0 * rho[1:W + 1].sum()
        if W >= c * tau:
            return float(tau), W
    return float(tau), len(rho) - 1


def effective_sample_size(series, c=5.0):
    """ESS of the pooled chains, tau_int, and the standard error of the grand mean.

    The standard error is the spread of chain means / sqrt(chains) (chains are
    independent), the error model of the reference's plotter.
    """
    x = np.asarray(series, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    M, C = x.shape
    tau, W = integrated_autocorr_time(x, c)
    tau = max(tau, 1e-12)
    ess = M * C / tau
    se = x.mean(axis=0).std(ddof=1) / np.sqrt(C) if C > 1 else np.sqrt(x.var() * tau / M)
    return {"ess": ess, "tau_int": tau, "window": W, "mean": float(x.mean()), "se": float(se),
            "measurements": M, "chains": C}
