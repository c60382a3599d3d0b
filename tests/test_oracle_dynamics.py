"""Pins for oracle.time_integrator and oracle.dynamics (CPU).

Reference: Time_Integrator.jl, Barotropic_Dynamics.jl,
Shallow_Water_Dynamics.jl, Dyn_Data.jl.  Independent checks: closed-form
single-mode recurrences, the Rossby-Haurwitz wave-4 exact solution of the
nondivergent barotropic vorticity equation, a balanced solid-body-rotation
shallow-water state that must stay steady, and the 2x2 implicit system.
"""
import numpy as np
import pytest

from oracle import FilteredLeapfrog, SpectralSphericalMesh
from oracle import dynamics as dyn

A, OMEGA = 6.371e6, 7.292e-5


def test_damping_table_and_implicit_damping():
    M = SpectralSphericalMesh(21, 22, 32, 64, 1, A)
    I = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.0, 1800, True, 0, 10)
    # normalised at degree N=21 (Time_Integrator.jl:31)
    assert I.damping[0, 21] == pytest.approx(1e-4, rel=1e-15)
    assert I.damping[3, 10] == pytest.approx(1e-4 * (10 * 11 / (21 * 22)) ** 4, rel=1e-14)
    # dQ = 0: backward-Euler damping Q -> Q/(1 + k Δt d)
    Qc = np.full((1, 23, 22), 1.0 + 2.0j)
    Qp = np.full((1, 23, 22), 3.0 - 1.0j)
    d = I.damping.T[None]
    dQ = np.zeros_like(Qc)
    I.compute_spectral_damping(Qc, Qp, dQ)
    Qn = Qc + 1800 * dQ
    np.testing.assert_allclose(Qn, Qc / (1 + 1800 * d), rtol=1e-14)
    I.update_init_step()
    dQ = np.zeros_like(Qc)
    I.compute_spectral_damping(Qc, Qp, dQ)
    np.testing.assert_allclose(Qp + 3600 * dQ, Qp / (1 + 3600 * d), rtol=1e-14)
    with pytest.raises(AssertionError):
        I.update_init_step()


def test_leapfrog_oscillator_and_asselin_filter():
    """dQ/dt = iωQ: unfiltered leapfrog keeps |Q| (physical mode |A|=1) and
    the phase error is O((ωΔt)^3) per step; with the filter the first step
    is forward Euler and the RA update matches Qc + r(Qp - 2Qc + Qn)."""
    w, dt = 1e-4, 100
    eig = -np.ones((2, 3))
    I = FilteredLeapfrog(0.0, 2, 0.0, eig, 0.0, dt, True, 0, 10)
    Qp = np.ones((1, 3, 2), complex)
    Qc = Qp.copy()
    Qn = np.zeros_like(Qc)
    I.filtered_leapfrog(1j * w * Qc, Qp, Qc, Qn)
    np.testing.assert_allclose(Qn, 1 + 1j * w * dt)
    I.update_init_step()
    Qp, Qc = Qc.copy(), Qn.copy()
    for _ in range(1000):
        Qn = np.zeros_like(Qc)
        I.filtered_leapfrog(1j * w * Qc, Qp, Qc, Qn)
        Qp, Qc = Qc, Qn
    exact = np.exp(1j * w * dt * 1001)
    assert np.abs(Qc - exact).max() < 1001 * (w * dt) ** 3
    J = FilteredLeapfrog(0.04, 2, 0.0, eig, 0.0, dt, False, 0, 10)
    qp, qc, dq = 0.3 + 0.1j, 0.5 - 0.2j, 0.01j
    a, b, c = (np.full((1, 3, 2), x, dtype=complex) for x in (qp, qc, 0))
    J.filtered_leapfrog(np.full((1, 3, 2), dq), a, b, c)
    qn = qp + 2 * dt * dq
    np.testing.assert_allclose(c, qn)
    np.testing.assert_allclose(b, qc + 0.04 * (qp - 2 * qc + qn))


def test_sw_implicit_correction_solves_2x2_system():
    M = SpectralSphericalMesh(21, 22, 32, 64, 1, A)
    I = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.5, 600, False, 0, 10)
    rng = np.random.default_rng(2)
    rnd = lambda: rng.standard_normal((1, 23, 22)) + 1j * rng.standard_normal((1, 23, 22))
    dc, dp, hc, hp, dD, dh = (rnd() for _ in range(6))
    h0 = 9.8e4
    D1, H1 = dD.copy(), dh.copy()
    dyn.sw_implicit_correction(I, dc, dp, hc, hp, h0, D1, H1)
    eig = M.laplacian_eig.T[None]
    mu = 0.5 * 1200
    # δD' = δD + eig (hc-hp) - μ eig δh',  δh' = δh + h0 (Dc-Dp) - μ h0 δD'
    r1 = D1 - (dD + eig * (hc - hp) - mu * eig * H1)
    r2 = H1 - (dh + h0 * (dc - dp) - mu * h0 * D1)
    assert np.abs(r1).max() < 1e-12 * np.abs(D1).max()
    assert np.abs(r2).max() < 1e-12 * np.abs(H1).max()


def _rh4_vorticity(M, omega_rh, K, lam_shift=0.0):
    mu = M.sin_theta[:, None]
    c = M.cos_theta[:, None]
    lam = M.lam_c[None, :] - lam_shift
    return (2 * omega_rh * mu - K * mu * c ** 4 * 30 * np.cos(4 * lam))[None]


def test_barotropic_rossby_haurwitz_wave4():
    """RH4 travels eastward at ν = (R(3+R)ω - 2Ω)/((1+R)(2+R)), R=4, as an
    exact solution; T42, Δt=1800 s, no filter / damping, 1 day."""
    N, nlat = 42, 64
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    w = K = 7.848e-6
    atmo = dyn.AtmoData("b", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    I = FilteredLeapfrog(0.0, 4, 0.0, M.laplacian_eig, 0.0, 1800, True, 0, 86400)
    d = dyn.DynData("b", N, N + 1, 2 * nlat, nlat, 1)
    z0 = _rh4_vorticity(M, w, K)
    M.trans_grid_to_spherical(z0, d.spe_vor_c)
    M.trans_spherical_to_grid(d.spe_vor_c, d.grid_vor)
    M.uv_grid_from_vor_div(d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)
    d.spe_vor_p[...] = d.spe_vor_c
    steps = 48
    dyn.barotropic_dynamics(M, atmo, d, I)
    I.update_init_step()
    for _ in range(steps - 1):
        dyn.barotropic_dynamics(M, atmo, d, I)
    R = 4
    nu = (R * (3 + R) * w - 2 * OMEGA) / ((1 + R) * (2 + R))
    z = _rh4_vorticity(M, w, K, lam_shift=nu * steps * 1800)
    wave = z - _rh4_vorticity(M, w, 0.0)
    err = np.abs(d.grid_vor - z).max() / np.abs(wave).max()
    assert err < 2e-3, err
    # the zonal part is steady to round-off
    zon = d.spe_vor_c[0, :, 0]
    zon0 = np.zeros_like(zon)
    s0 = M.spec_zeros(1)
    M.trans_grid_to_spherical(z0, s0)
    np.testing.assert_allclose(zon, s0[0, :, 0], atol=1e-12 * np.abs(zon).max())


def test_shallow_water_balanced_state_is_steady():
    """Solid-body zonal flow u = u0 cos φ with geopotential
    h = h0 - (aΩu0 + u0²/2) sin²φ is an exact steady state."""
    N, nlat = 21, 32
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = dyn.AtmoData("sw", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    I = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.5, 600, True, 0, 86400)
    d = dyn.DynData("sw", N, N + 1, 2 * nlat, nlat, 1)
    u0, h0 = 20.0, 9.80616 * 1e4
    mu = M.sin_theta[None, :, None]
    u = u0 * np.sqrt(1 - mu ** 2) * np.ones((1, nlat, 2 * nlat))
    h = h0 - (A * OMEGA * u0 + 0.5 * u0 ** 2) * mu ** 2 * np.ones((1, nlat, 2 * nlat))
    M.vor_div_from_grid_uv(u, np.zeros_like(u), d.spe_vor_c, d.spe_div_c)
    M.trans_grid_to_spherical(h, d.spe_lnps_c)
    M.trans_spherical_to_grid(d.spe_vor_c, d.grid_vor)
    M.trans_spherical_to_grid(d.spe_lnps_c, d.grid_ps_c)
    M.uv_grid_from_vor_div(d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    for f in ("vor", "div", "lnps"):
        getattr(d, f"spe_{f}_p")[...] = getattr(d, f"spe_{f}_c")
        getattr(d, f"spe_{f}_n")[...] = getattr(d, f"spe_{f}_c")
    vor0, h0s = d.spe_vor_c.copy(), d.spe_lnps_c.copy()
    for i in range(12):
        dyn.shallow_water_physics(d, 0.0, 0.0, np.zeros((1, nlat, 2 * nlat)))
        dyn.shallow_water_dynamics(M, atmo, h0, d, I)
        if i == 0:
            I.update_init_step()
    # the ∇⁸ damping of degrees 1-2 (d ~ 1e-4 (2/462)^4 s⁻¹ over 4 h) is ~1e-9
    np.testing.assert_allclose(d.spe_vor_c, vor0, atol=5e-9 * np.abs(vor0).max())
    np.testing.assert_allclose(d.spe_lnps_c, h0s, atol=5e-9 * np.abs(h0s).max())
    assert np.abs(d.spe_div_c).max() < 2e-8 * np.abs(vor0).max()
