"""Pins for oracle.mesh (CPU): transforms and spectral operators.

Reference: src/Atmos_Spectral/Spectral_Spherical_Mesh.jl (SSM).  The
reference's own checks are norm prints (SSM:763-1008); here they become
assertions against closed forms built from scipy's Legendre functions
(tests/harmonics.py), so a dropped term, wrong sign or transposed index in
the oracle fails.
"""
import numpy as np
import pytest

from oracle import SpectralSphericalMesh
from tests.harmonics import (harmonic_grid, harmonic_lambda_derivative,
                             harmonic_mu_derivative)

RADIUS = 6371.0e3
MESHES = {}


def mesh(N, nlat, nd=1, radius=RADIUS):
    key = (N, nlat, nd, radius)
    if key not in MESHES:
        MESHES[key] = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, nd, radius)
    return MESHES[key]


CASES = [(0, 0, 1.3), (0, 5, -0.7), (1, 1, 0.4 - 0.9j), (3, 8, 1.1 + 0.2j),
         (7, 7, -0.5 + 0.5j), (11, 20, 0.25j), (20, 21, 0.8 - 0.1j), (21, 21, 1.0 + 1.0j)]


@pytest.mark.parametrize("m,n,c", CASES)
def test_forward_single_harmonic(m, n, c):
    M = mesh(21, 32)
    g = harmonic_grid(m, n, c, M.sin_theta, M.lam_c)[None]
    s = M.trans_grid_to_spherical(g, M.spec_zeros())
    expect = M.spec_zeros()
    expect[0, n, m] = c if m > 0 else np.real(c)
    np.testing.assert_allclose(s, expect, atol=1e-13)


@pytest.mark.parametrize("m,n,c", CASES + [(5, 22, 0.3 + 0.7j), (0, 22, 2.0)])
def test_inverse_single_harmonic(m, n, c):
    """Degree N+1 (n=22 at T21) is summed by the inverse (SSM:196)."""
    M = mesh(21, 32)
    s = M.spec_zeros()
    s[0, n, m] = c if m > 0 else np.real(c)
    g = M.trans_spherical_to_grid(s, M.grid_zeros(1))
    np.testing.assert_allclose(g[0], harmonic_grid(m, n, c, M.sin_theta, M.lam_c), atol=1e-12)


def test_forward_zeroes_top_degree_and_keeps_lower_triangle():
    M = mesh(21, 32)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((1, 32, 64))
    s = np.full(M.spec_zeros().shape, 7.0 + 7.0j)
    M.trans_grid_to_spherical(g, s)
    assert np.all(s[0, 22, :] == 0)
    for m in range(22):
        for n in range(m):
            if n != 22:
                assert s[0, n, m] == 7.0 + 7.0j


def test_round_trip_idempotent_t31():
    """SSM:763-787 (test_fourier) as an assertion, seeded."""
    M = SpectralSphericalMesh(31, 32, 48, 96, 2, 1.0)
    rng = np.random.default_rng(7)
    p0 = rng.random((2, 48, 96))
    s0 = M.trans_grid_to_spherical(p0, M.spec_zeros())
    p1 = M.trans_spherical_to_grid(s0, M.grid_zeros())
    s1 = M.trans_grid_to_spherical(p1, M.spec_zeros())
    p2 = M.trans_spherical_to_grid(s1, M.grid_zeros())
    np.testing.assert_allclose(s1, s0, atol=1e-13)
    np.testing.assert_allclose(p2, p1, atol=1e-13)


def _psi_fields(M, m, n, c):
    """u, v on the grid for stream function psi = Y(m,n,c):
    u = -(1/r) dpsi/dtheta = -(cos/r) dpsi/dmu, v = (1/(r cos)) dpsi/dlambda."""
    r, mu, lam = M.radius, M.sin_theta, M.lam_c
    cos = M.cos_theta[:, None]
    u = -(cos / r) * harmonic_mu_derivative(m, n, c, mu, lam)
    v = harmonic_lambda_derivative(m, n, c, mu, lam) / (r * cos)
    return u, v


def _chi_fields(M, m, n, c):
    """u, v for velocity potential chi = Y(m,n,c): u = (1/(r cos)) dchi/dlambda,
    v = (1/r) dchi/dtheta = (cos/r) dchi/dmu."""
    r, mu, lam = M.radius, M.sin_theta, M.lam_c
    cos = M.cos_theta[:, None]
    u = harmonic_lambda_derivative(m, n, c, mu, lam) / (r * cos)
    v = (cos / r) * harmonic_mu_derivative(m, n, c, mu, lam)
    return u, v


@pytest.mark.parametrize("m,n,c", [(0, 1, 1.0), (0, 4, -2.0), (2, 3, 1 - 1j), (5, 12, 0.3 + 0.8j), (21, 21, 1j)])
def test_vor_div_and_uv_from_stream_and_potential(m, n, c):
    M = mesh(21, 32)
    cc = c if m > 0 else np.real(c)
    lap = -n * (n + 1) / M.radius ** 2
    for kind in ("psi", "chi"):
        u, v = (_psi_fields if kind == "psi" else _chi_fields)(M, m, n, c)
        vor, div = M.spec_zeros(), M.spec_zeros()
        M.vor_div_from_grid_uv(u[None], v[None], vor, div)
        ev, ed = M.spec_zeros(), M.spec_zeros()
        (ev if kind == "psi" else ed)[0, n, m] = lap * cc
        scale = abs(lap)
        np.testing.assert_allclose(vor / scale, ev / scale, atol=1e-12)
        np.testing.assert_allclose(div / scale, ed / scale, atol=1e-12)
        gu, gv = M.grid_zeros(1), M.grid_zeros(1)
        M.uv_grid_from_vor_div(ev, ed, gu, gv)
        uscale = np.abs(u).max() + np.abs(v).max()
        np.testing.assert_allclose(gu[0] / uscale, u / uscale, atol=1e-12)
        np.testing.assert_allclose(gv[0] / uscale, v / uscale, atol=1e-12)


def test_rigid_rotation_vorticity():
    """SSM:821-840 closed form (the one exact case of velocity_profile), T85, radius 5."""
    M = SpectralSphericalMesh(85, 86, 128, 256, 1, 5.0)
    U, beta = 2 * np.pi / 256.0, np.pi / 2.0
    cos, sin, lam = M.cos_theta[:, None], M.sin_theta[:, None], M.lam_c[None, :]
    u = U * (np.cos(beta) * cos + np.sin(beta) * np.cos(lam) * sin)
    v = -U * np.sin(beta) * np.sin(lam) * np.ones_like(cos)
    vor_g = 2 * U / 5.0 * (np.cos(beta) * sin - np.sin(beta) * np.cos(lam) * cos)
    vor, div = M.spec_zeros(), M.spec_zeros()
    M.vor_div_from_grid_uv(u[None], v[None], vor, div)
    gv = M.trans_spherical_to_grid(vor, M.grid_zeros(1))
    gd = M.trans_spherical_to_grid(div, M.grid_zeros(1))
    np.testing.assert_allclose(gv[0], vor_g, atol=1e-13)
    np.testing.assert_allclose(gd[0], 0.0, atol=1e-13)
    gu, gvv = M.grid_zeros(1), M.grid_zeros(1)
    M.uv_grid_from_vor_div(vor, div, gu, gvv)
    np.testing.assert_allclose(gu[0], u, atol=1e-13)
    np.testing.assert_allclose(gvv[0], v, atol=1e-13)


@pytest.mark.parametrize("m,n,c", [(0, 3, 1.5), (4, 9, 0.2 - 0.6j), (20, 20, 1 + 0j)])
def test_gradients_and_advection(m, n, c):
    M = mesh(21, 32)
    r, mu, lam = M.radius, M.sin_theta, M.lam_c
    cos = M.cos_theta[:, None]
    s = M.spec_zeros()
    s[0, n, m] = c if m > 0 else np.real(c)
    gl, gt = M.grid_zeros(1), M.grid_zeros(1)
    M.compute_gradients(s, gl, gt)
    el = harmonic_lambda_derivative(m, n, c, mu, lam) / (r * cos)
    et = cos * harmonic_mu_derivative(m, n, c, mu, lam) / r
    sc = np.abs(el).max() + np.abs(et).max()
    np.testing.assert_allclose(gl[0] / sc, el / sc, atol=1e-12)
    np.testing.assert_allclose(gt[0] / sc, et / sc, atol=1e-12)
    rng = np.random.default_rng(3)
    u, v, dh = (rng.standard_normal((1, 32, 64)) for _ in range(3))
    out = dh.copy()
    M.add_horizontal_advection(s, u, v, out)
    np.testing.assert_allclose(out, dh - (u * el + v * et), atol=1e-12 * sc * 10)


def test_laplacian_and_area_mean():
    M = mesh(21, 32)
    s = np.ones(M.spec_zeros().shape, np.complex128)
    M.apply_laplacian(s)
    for mm in range(22):
        for nn in range(23):
            expect = -nn * (nn + 1) / M.radius ** 2 if nn >= mm else 0.0
            assert s[0, nn, mm] == pytest.approx(expect, rel=1e-15, abs=0)
    assert M.area_weighted_global_mean(np.ones((1, 32, 64))) == pytest.approx(1.0, abs=1e-15)
    g = harmonic_grid(3, 5, 1 + 1j, M.sin_theta, M.lam_c)[None] + 0.25
    assert M.area_weighted_global_mean(g) == pytest.approx(0.25, abs=1e-14)
    g2 = harmonic_grid(0, 2, 1.0, M.sin_theta, M.lam_c)[None]
    assert M.area_weighted_global_mean(g2) == pytest.approx(0.0, abs=1e-14)
