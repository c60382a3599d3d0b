"""Closed-form spherical-harmonic fields used by the parity tests.

Independent of the oracle: P_n^m comes from scipy's ``lpmv`` (Condon-Shortley
phase removed, unit normalisation 1/2 int P^2 dmu = 1), derivatives from the
textbook recurrence.  Conventions follow the reference's expansion
(Spectral_Spherical_Mesh.jl:145-160): a coefficient c at (m, n) with m > 0
contributes P_n^m(mu) * 2 Re(c e^{i m lambda}); at m = 0 it contributes
c P_n^0 (c real).
"""
import numpy as np
from scipy.special import gammaln, lpmv


def pbar(m, n, mu):
    if n < m:
        return np.zeros_like(mu)
    norm = np.exp(0.5 * (np.log(2 * n + 1) + gammaln(n - m + 1) - gammaln(n + m + 1)))
    return (-1) ** m * norm * lpmv(m, n, mu)


def dpbar(m, n, mu):
    """dP/dmu via (1-mu^2) P' = (n+1) mu P_n - (n-m+1) P_{n+1} (scaled)."""
    norm = np.exp(0.5 * (np.log(2 * n + 1) + gammaln(n - m + 1) - gammaln(n + m + 1)))
    p = (-1) ** m * norm * lpmv(m, n, mu)
    p1 = (-1) ** m * norm * lpmv(m, n + 1, mu)
    return ((n + 1) * mu * p - (n - m + 1) * p1) / (1 - mu ** 2)


def harmonic_grid(m, n, c, mu, lam):
    """Grid values [nlat, nlon] of the single harmonic (m, n, c)."""
    p = pbar(m, n, mu)[:, None]
    if m == 0:
        return np.real(c) * p * np.ones_like(lam)[None, :]
    return p * 2.0 * np.real(c * np.exp(1j * m * lam))[None, :]


def harmonic_lambda_derivative(m, n, c, mu, lam):
    """d/dlambda of harmonic_grid."""
    p = pbar(m, n, mu)[:, None]
    return p * 2.0 * np.real(1j * m * c * np.exp(1j * m * lam))[None, :]


def harmonic_mu_derivative(m, n, c, mu, lam):
    """d/dmu of harmonic_grid."""
    dp = dpbar(m, n, mu)[:, None]
    if m == 0:
        return np.real(c) * dp * np.ones_like(lam)[None, :]
    return dp * 2.0 * np.real(c * np.exp(1j * m * lam))[None, :]
