"""Host-side profiles of the initial-condition builders, pinned by what the
mathematics fixes: the jet's peak and support, gradient balance of the
geopotential (central differences of the quadrature result against the
balance equation), convergence of the quadrature, the prescribed global mean,
and the bump's amplitude, position and longitude wrap; the Rossby-Haurwitz
vorticity against the Laplacian of its streamfunction (spectral identity
∇²Y_n = -n(n+1)/a² Y_n, applied with numpy on the grid by finite differences)."""
import numpy as np
import pytest

import __graft_entry__

__graft_entry__.load_package()
from jgcm_amd import initial_conditions as ic  # noqa: E402

A, OMEGA, G = 6.371e6, 7.292e-5, 9.80616
P = ic.GALEWSKY


def gauss_grid(nlat):
    x, w = np.polynomial.legendre.leggauss(nlat)
    return np.arcsin(x), w


def test_jet_peak_and_support():
    lat = np.linspace(-np.pi / 2, np.pi / 2, 20001)
    u = ic.galewsky_zonal_wind(lat)
    assert u.max() == pytest.approx(80.0, rel=1e-6)
    assert abs(lat[np.argmax(u)] - np.pi / 4) < 2e-4          # peak at (φ0 + φ1) / 2
    assert ic.galewsky_zonal_wind(np.array([np.pi / 4]))[0] == pytest.approx(80.0, rel=1e-15)
    out = (lat <= P["lat0"]) | (lat >= P["lat1"])
    assert np.all(u[out] == 0.0) and np.all(u >= 0.0)
    # positive wherever exp(1/((φ-φ0)(φ-φ1))) does not underflow
    away = (lat > P["lat0"] + 0.01) & (lat < P["lat1"] - 0.01)
    assert np.all(u[away] > 0.0)


def test_geopotential_gradient_balance():
    h = 1e-5
    probe = np.linspace(P["lat0"] + 0.02, P["lat1"] - 0.02, 17)
    phi_p = ic.galewsky_geopotential_drop(probe + h, A, OMEGA)
    phi_m = ic.galewsky_geopotential_drop(probe - h, A, OMEGA)
    u = ic.galewsky_zonal_wind(probe)
    rhs = -A * u * (2 * OMEGA * np.sin(probe) + u * np.tan(probe) / A)
    fd = (phi_p - phi_m) / (2 * h)
    np.testing.assert_allclose(fd, rhs, rtol=1e-6, atol=1e-6 * np.abs(rhs).max())


def test_geopotential_flat_outside_jet_and_mean():
    lat, w = gauss_grid(128)
    phi = ic.galewsky_geopotential(lat, w, A, OMEGA, G, 1e4)
    assert np.sum(w * phi) / np.sum(w) == pytest.approx(G * 1e4, rel=1e-14)
    south, north = lat <= P["lat0"], lat >= P["lat1"]
    assert np.ptp(phi[south]) < 1e-9 * G * 1e4 and np.ptp(phi[north]) < 1e-9 * G * 1e4
    # the jet (westerly, northern hemisphere) lowers the height to its north
    assert phi[north][0] < phi[south][0]


def test_geopotential_quadrature_converged():
    lat, w = gauss_grid(128)
    a = ic.galewsky_geopotential(lat, w, A, OMEGA, G, 1e4)
    b = ic.galewsky_geopotential(lat, w, A, OMEGA, G, 1e4, panels=256, order=32)
    drop = np.ptp(b)
    assert np.abs(a - b).max() < 1e-11 * drop


def test_bump_amplitude_position_and_wrap():
    lat = np.array([np.pi / 4, 0.0, -np.pi / 4])
    lon_c = 6.2
    lon = np.array([lon_c, lon_c - 0.3, (lon_c + 0.3) - 2 * np.pi])
    b = ic.galewsky_bump(lat, lon, lon_c)
    assert b[0, 0] == pytest.approx(120.0 * np.cos(np.pi / 4), rel=1e-15)
    # symmetric about λ_c across the 2π wrap
    assert b[0, 1] == pytest.approx(b[0, 2], rel=1e-12)
    assert b[0, 1] == pytest.approx(120.0 * np.cos(np.pi / 4) * np.exp(-(0.3 * 3) ** 2), rel=1e-12)
    assert b[1, 0] < 1e-40 and b[2, 0] < 1e-40


def test_rossby_haurwitz_is_laplacian_of_streamfunction():
    # ψ = -a²ω sinφ + a²K cos^4 φ sinφ cos 4λ; ζ = ∇²ψ by second-order differences
    om = K = 7.848e-6
    lat = np.linspace(-1.2, 1.2, 241)
    lon = np.linspace(0, 2 * np.pi, 361)
    dl, dp = lon[1] - lon[0], lat[1] - lat[0]
    L, Lm = np.meshgrid(lat, lon, indexing="ij")
    psi = A * A * (-om * np.sin(L) + K * np.cos(L) ** 4 * np.sin(L) * np.cos(4 * Lm))
    c = np.cos(L)
    d2l = (np.roll(psi, -1, 1) - 2 * psi + np.roll(psi, 1, 1)) / dl ** 2
    flux = c * np.gradient(psi, dp, axis=0)
    lap = (d2l / c ** 2 + np.gradient(flux, dp, axis=0) / c) / A ** 2
    z = ic.rossby_haurwitz_vorticity(lat, lon, om, K)
    inner = (slice(5, -5), slice(1, -1))
    err = np.abs(lap[inner] - z[inner]).max() / np.abs(z).max()
    assert err < 2e-3


def test_spectral_noise_layout():
    rng = np.random.default_rng(0)
    s = ic.spectral_noise(5, 2, 1e-8, rng)
    assert s.shape == (2, 7, 6)
    assert np.all(s[:, :, 0].imag == 0)
    assert np.all(s[:, 6, :] == 0)
    for m in range(6):
        assert np.all(s[:, :m, m] == 0)
    assert 0.5e-8 < np.std(s[:, 3:6, 1].real) < 2e-8
