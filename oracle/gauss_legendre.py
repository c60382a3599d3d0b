"""Gaussian quadrature and normalised associated Legendre functions (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/Atmos_Spectral/Gauss_And_Legendre.jl``:

* ``compute_gaussian``  <- ``Compute_Gaussian!``  (Gauss_And_Legendre.jl:69-118)
* ``compute_legendre``  <- ``Compute_Legendre!``  (Gauss_And_Legendre.jl:4-66)

Pinned by ``tests/test_oracle_gauss_legendre.py``: numpy's Golub-Welsch
``leggauss`` nodes/weights, the closed forms of Gauss_And_Legendre.jl:132-143,
the orthonormality assertion of Gauss_And_Legendre.jl:162-168 and scipy's
``lpmv`` with the Condon-Shortley phase removed.
"""
import numpy as np


def compute_gaussian(n):
    """Gauss-Legendre nodes (south to north) and weights.

    Newton iteration per node exactly as Gauss_And_Legendre.jl:81-115:
    initial guess cos(pi (i - 1/4)/(n + 1/2)), tolerance 1e-15, the weight
    uses the derivative from the last *pre-update* iterate (:102-114).
    Vectorised over the n/2 nodes; a node stops iterating on the first
    iteration whose update is <= tol, as the scalar loop's ``break`` does.
    """
    itermax = 10000
    tol = 1.0e-15
    n_half = n // 2
    i = np.arange(1, n_half + 1, dtype=np.float64)
    z = np.cos(np.pi * (i - 0.25) / (n + 0.5))
    dp = np.zeros(n_half)
    active = np.ones(n_half, dtype=bool)
    for _ in range(itermax):
        za = z[active]
        p2 = np.zeros_like(za)
        p1 = np.ones_like(za)
        for j in range(1, n + 1):
            p3 = p2
            p2 = p1
            p1 = ((2.0 * j - 1.0) * za * p2 - (j - 1.0) * p3) / j
        dpa = n * (za * p1 - p2) / (za * za - 1.0)
        z1 = za
        znew = z1 - p1 / dpa
        idx = np.nonzero(active)[0]
        dp[idx] = dpa
        z[idx] = znew
        done = np.abs(znew - z1) <= tol
        active[idx[done]] = False
        if not active.any():
            break
    sin_theta = np.zeros(n)
    wts = np.zeros(n)
    sin_theta[:n_half] = -z
    sin_theta[n - 1:n - n_half - 1:-1] = z
    w = 2.0 / ((1.0 - z * z) * dp * dp)
    wts[:n_half] = w
    wts[n - 1:n - n_half - 1:-1] = w
    return sin_theta, wts


def compute_legendre(num_fourier, num_spherical, sin_theta, nlat):
    """Normalised P_n^m(mu) and dP_n^m/dmu, no Condon-Shortley phase.

    Returns ``qnm[m, n, j]`` for degrees n = 0..num_spherical (the table is
    built to num_spherical+1 and the top degree dropped, :21 and :64) and
    ``dqnm[m, n, j]`` for n = 0..num_spherical.  Normalisation
    1/2 int P^2 dmu = 1 (Gauss_And_Legendre.jl:5-14).
    """
    N, ns = num_fourier, num_spherical
    s = np.asarray(sin_theta, dtype=np.float64)
    cos_theta = np.sqrt(1.0 - s ** 2)
    qnm = np.zeros((N + 1, ns + 2, nlat))
    dqnm = np.zeros((N + 1, ns + 1, nlat))
    eps = np.zeros((N + 1, ns + 2))

    qnm[0, 0, :] = 1.0
    for m in range(1, N + 1):                       # :30-32
        qnm[m, m, :] = np.sqrt((2 * m + 1) / (2 * m)) * cos_theta * qnm[m - 1, m - 1, :]
    for m in range(0, N + 1):                       # :34-36 (Julia m-1 -> m)
        qnm[m, m + 1, :] = np.sqrt(2 * m + 3) * s * qnm[m, m, :]
    for m in range(0, N + 1):                       # :38-43
        for l in range(m, ns + 2):
            eps[m, l] = np.sqrt((l * l - m * m) / (4.0 * l * l - 1.0))
    for m in range(0, N + 1):                       # :45-51
        for l in range(m + 2, ns + 2):
            qnm[m, l, :] = (s * qnm[m, l - 1, :] - eps[m, l - 1] * qnm[m, l - 2, :]) / eps[m, l]
    c2 = cos_theta ** 2
    for m in range(0, N + 1):                       # :53-62
        for l in range(m, ns + 1):
            if l == m:
                dqnm[m, l, :] = (-l * eps[m, l + 1] * qnm[m, l + 1, :]) / c2
            else:
                dqnm[m, l, :] = (-l * eps[m, l + 1] * qnm[m, l + 1, :]
                                 + (l + 1) * eps[m, l] * qnm[m, l - 1, :]) / c2
    return qnm[:, :ns + 1, :].copy(), dqnm
