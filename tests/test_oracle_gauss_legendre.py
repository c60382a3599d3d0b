"""Pins for oracle.gauss_legendre (CPU).

Reference: src/Atmos_Spectral/Gauss_And_Legendre.jl.  The reference's own
checks are the closed forms at :132-143 and the orthonormality assertion at
:162-168 (made two-sided here).  Independent pins: numpy's Golub-Welsch
Gauss-Legendre rule and scipy's associated Legendre functions.
"""
import numpy as np
import pytest
from scipy.special import lpmv, gammaln

from oracle import compute_gaussian, compute_legendre


@pytest.mark.parametrize("n", [32, 64, 128])
def test_gaussian_matches_golub_welsch(n):
    x, w = np.polynomial.legendre.leggauss(n)          # ascending nodes
    s, wts = compute_gaussian(n)
    assert np.all(np.diff(s) > 0), "nodes must run south -> north (Gauss_And_Legendre.jl:113)"
    np.testing.assert_allclose(s, x, rtol=0, atol=2e-15)
    # numpy's weights are only good to ~1e-12; the Newton weights are better
    np.testing.assert_allclose(wts, w, rtol=0, atol=5e-14)
    assert abs(wts.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("n", [32, 64, 128])
def test_gaussian_weights_high_precision(n):
    import mpmath as mp
    mp.mp.dps = 40
    s, wts = compute_gaussian(n)
    for i in range(n // 2):
        z = mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(s[i]))
        dp = mp.diff(lambda t: mp.legendre(n, t), z)
        wt = 2 / ((1 - z ** 2) * dp ** 2)
        assert abs(s[i] - float(z)) <= 2e-16 * max(1.0, abs(float(z))) * 4
        assert abs(wts[i] - float(wt)) <= 1e-12 * float(wt)


def test_legendre_closed_forms_t21():
    """Gauss_And_Legendre.jl:120-156 (T21, nθ=32)."""
    N, nlat = 21, 32
    s, w = compute_gaussian(nlat)
    q, dq = compute_legendre(N, N + 1, s, nlat)
    c = np.sqrt(1 - s ** 2)
    checks = {
        (3, 3): (np.sqrt(35) / 4 * c ** 3, np.sqrt(35) / 4 * c * 1.5 * (-2 * s)),
        (2, 3): (np.sqrt(105 / 8) * (s - s ** 3), np.sqrt(105 / 8) * (1 - 3 * s ** 2)),
        (0, 2): (np.sqrt(5) / 2 * (3 * s ** 2 - 1), np.sqrt(5) / 2 * 6 * s),
        (0, 3): (np.sqrt(7) / 2 * (5 * s ** 3 - 3 * s), np.sqrt(7) / 2 * (15 * s ** 2 - 3)),
        (1, 3): (np.sqrt(21) / 4 * (5 * s ** 2 - 1) * c,
                 np.sqrt(21) / 4 * 10 * s * c + np.sqrt(21) / 4 * (5 * s ** 2 - 1) / c * 0.5 * (-2 * s)),
    }
    for (m, n), (p, dp) in checks.items():
        np.testing.assert_allclose(q[m, n], p, atol=1e-13)
        np.testing.assert_allclose(dq[m, n], dp, atol=1e-12)


@pytest.mark.parametrize("N,nlat", [(21, 32), (42, 64)])
def test_legendre_against_scipy(N, nlat):
    """Normalised P_n^m = sqrt((2n+1)(n-m)!/(n+m)!) P_n^m without the
    Condon-Shortley phase that scipy's lpmv includes."""
    s, _ = compute_gaussian(nlat)
    q, dq = compute_legendre(N, N + 1, s, nlat)
    assert q.shape == (N + 1, N + 2, nlat) and dq.shape == (N + 1, N + 2, nlat)
    for m in range(N + 1):
        for n in range(m, N + 2):
            norm = np.exp(0.5 * (np.log(2 * n + 1) + gammaln(n - m + 1) - gammaln(n + m + 1)))
            ref = (-1) ** m * norm * lpmv(m, n, s)
            np.testing.assert_allclose(q[m, n], ref, rtol=1e-11, atol=1e-11)
            # dP/dμ from the standard recurrence (1-μ²) P'_n^m = (n+1) μ P_n^m - (n-m+1) P_{n+1}^m
            p1 = (-1) ** m * lpmv(m, n + 1, s) * norm
            dref = ((n + 1) * s * ref - (n - m + 1) * p1) / (1 - s ** 2)
            np.testing.assert_allclose(dq[m, n], dref, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(dref).max()))
        for n in range(m):
            assert not q[m, n].any()


@pytest.mark.parametrize("N,nlat", [(21, 32), (85, 128)])
def test_legendre_orthonormal(N, nlat):
    """Gauss_And_Legendre.jl:158-168 with |.| (the reference's assert is one-sided)."""
    s, w = compute_gaussian(nlat)
    q, _ = compute_legendre(N, N + 1, s, nlat)
    for m in range(N + 1):
        G = 0.5 * (q[m, m:] * w) @ q[m, m:].T
        np.testing.assert_allclose(G, np.eye(G.shape[0]), atol=1e-12)
