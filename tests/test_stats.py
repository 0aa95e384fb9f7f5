"""Autocorrelation / ESS estimators against processes with known answers (CPU)."""
import numpy as np

from feynsimul_amd.stats import autocorrelation, integrated_autocorr_time, effective_sample_size


def _ar1(phi, M, C, rng):
    x = np.empty((M, C))
    x[0] = rng.normal(0, 1 / np.sqrt(1 - phi * phi), C)
    e = rng.normal(0, 1, (M, C))
    for t in range(1, M):
        x[t] = phi * x[t - 1] + e[t]
    return x


def test_ar1_tau_matches_closed_form():
    rng = np.random.default_rng(0)
    for phi in (0.0, 0.5, 0.9):
        x = _ar1(phi, 2000, 64, rng)
        tau, _ = integrated_autocorr_time(x)
        exact = (1 + phi) / (1 - phi)
        assert abs(tau / exact - 1) < 0.1, (phi, tau, exact)
        rho = autocorrelation(x)
        assert abs(rho[1] - phi) < 0.03


def test_ess_consistent_with_standard_error():
    rng = np.random.default_rng(1)
    phi = 0.8
    x = _ar1(phi, 400, 512, rng)
    r = effective_sample_size(x)
    # Var(grand mean) = Var(x) / ESS for an equilibrium stationary process
    se_from_ess = np.sqrt(x.var() / r["ess"])
    assert abs(se_from_ess / r["se"] - 1) < 0.15
    assert abs(r["ess"] / (400 * 512 * (1 - phi) / (1 + phi)) - 1) < 0.15
