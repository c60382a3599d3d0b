"""Initial-condition builders for the barotropic and shallow-water experiments.

The reference hard-codes its initial states in the experiment drivers
(exp/Barotropic/Barotropic.jl:52-66, exp/Shallow_Water/SW.jl:57-68): it sets
grid winds and heights on the host, then forms the spectral state with
Vor_Div_From_Grid_UV! / Trans_Grid_To_Spherical! and the grid state with
Trans_Spherical_To_Grid!.  The builders here take the same steps, in the same
order, through the library's transforms; only the analytic grid fields are
evaluated on the host (set-up, once per run).  The Held-Suarez state is
``primitive.Spectral_Initialize_Fields`` (Spectral_Dynamics.jl:372-449).

* ``Rossby_Haurwitz_Wave4`` — the wave-4 Rossby-Haurwitz vorticity
  (ω = K = 7.848e-6 s^-1) with seeded N(0, σ) noise on every valid spectral
  coefficient (the T42 barotropic case).
* ``Galewsky_Jet`` — the barotropically unstable mid-latitude jet of
  Galewsky, Scott & Polvani (2004, Tellus 56A, 429-440): zonal wind
  u(φ) = u_max/e_n · exp(1/((φ-φ0)(φ-φ1))) for φ0 < φ < φ1, geopotential in
  gradient balance (dΦ/dφ = -a u (2Ω sin φ + u tan φ / a), integrated by
  composite Gauss-Legendre quadrature), plus the Gaussian height bump
  ĥ cos φ exp(-(λ'/α)^2) exp(-((φ2-φ)/β)^2) at a seeded longitude.
  The shallow-water state carries the geopotential gh in the height slots
  (Shallow_Water_Dynamics.jl:109-115: spe_lnps_*, grid_ps_*), as the
  reference's δdiv -= ∇²(KE + h) requires.
"""
import numpy as np
import torch

from .mesh import (Trans_Grid_To_Spherical, Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div,
                   Vor_Div_From_Grid_UV)


def _host_to(dst, arr):
    host = np.array(arr, dtype=np.float64, order="C", copy=True)   # writable, contiguous
    dst.copy_(torch.from_numpy(host).reshape(dst.shape))


def _previous_equals_current(d, spectral, grid):
    """p <- c for the named fields: the first (forward) step's implicit
    correction and damping read the previous state and assume it equals the
    current one (Shallow_Water_Dynamics.jl:38-40)."""
    for f in spectral:
        getattr(d, f"spe_{f}_p").copy_(getattr(d, f"spe_{f}_c"))
    for f in grid:
        getattr(d, f"grid_{f}_p").copy_(getattr(d, f"grid_{f}_c"))


def spectral_noise(N, nlev, sigma, rng):
    """Complex N(0, σ) on every valid coefficient of a [nlev][N+2][N+1] array
    (m = 0 real, m > n zero, degree N+1 zero)."""
    s = sigma * (rng.standard_normal((nlev, N + 2, N + 1))
                 + 1j * rng.standard_normal((nlev, N + 2, N + 1)))
    s[:, :, 0] = s[:, :, 0].real
    for m in range(N + 1):
        s[:, :m, m] = 0
    s[:, N + 1, :] = 0
    return s


def rossby_haurwitz_vorticity(lat, lon, omega=7.848e-6, K=7.848e-6, R=4):
    """ζ(φ, λ) = 2ω sin φ - K sin φ cos^R φ (R²+3R+2) cos Rλ, the relative
    vorticity of the streamfunction ψ = -a²ω sin φ + a²K cos^R φ sin φ cos Rλ."""
    mu, c = np.sin(lat)[:, None], np.cos(lat)[:, None]
    return 2 * omega * mu - K * mu * c ** R * (R * R + 3 * R + 2) * np.cos(R * lon[None, :])


def Rossby_Haurwitz_Wave4(mesh, dyn_data, omega=7.848e-6, K=7.848e-6, noise=1.0e-8, seed=42):
    """Barotropic state: spectral vorticity of the wave-4 Rossby-Haurwitz
    field plus seeded noise of standard deviation ``noise`` (s^-1, real and
    imaginary parts) on the valid coefficients; grid vorticity and winds from
    it (exp/Barotropic/Barotropic.jl:52-66 order: spectral vorticity first,
    then Trans_Spherical_To_Grid! and UV_Grid_From_Vor_Div!)."""
    d = dyn_data
    z = rossby_haurwitz_vorticity(np.asarray(mesh.θc), np.asarray(mesh.λc), omega, K)
    grid = torch.empty_like(d.grid_vor)
    _host_to(grid, np.broadcast_to(z[None], d.grid_vor.shape))
    torch.cuda.synchronize()
    Trans_Grid_To_Spherical(mesh, grid, d.spe_vor_c)
    mesh.synchronize()
    if noise:
        rng = np.random.default_rng(seed)
        eps = spectral_noise(mesh.num_fourier, d.spe_vor_c.shape[0], noise, rng)
        d.spe_vor_c.add_(torch.from_numpy(eps).to(d.spe_vor_c.device))
        torch.cuda.synchronize()
    d.spe_div_c.zero_()
    torch.cuda.synchronize()
    Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    mesh.synchronize()
    _previous_equals_current(d, ("vor", "div"), ("u", "v"))
    torch.cuda.synchronize()


# Galewsky et al. (2004) constants: jet edges φ0 = π/7, φ1 = π/2 - φ0, bump at φ2 = π/4
GALEWSKY = dict(u_max=80.0, lat0=np.pi / 7, lat1=np.pi / 2 - np.pi / 7, lat2=np.pi / 4,
                bump=120.0, alpha=1.0 / 3.0, beta=1.0 / 15.0)


def galewsky_zonal_wind(lat, u_max=80.0, lat0=np.pi / 7, lat1=np.pi / 2 - np.pi / 7):
    """u(φ) = u_max / e_n · exp(1 / ((φ-φ0)(φ-φ1))) inside (φ0, φ1), zero
    outside; e_n = exp(-4 / (φ1-φ0)²) makes the maximum u_max at (φ0+φ1)/2."""
    lat = np.asarray(lat, dtype=np.float64)
    u = np.zeros_like(lat)
    inside = (lat > lat0) & (lat < lat1)
    e_n = np.exp(-4.0 / (lat1 - lat0) ** 2)
    x = lat[inside]
    u[inside] = u_max / e_n * np.exp(1.0 / ((x - lat0) * (x - lat1)))
    return u


def galewsky_geopotential_drop(lat, radius, omega, u_max=80.0, lat0=np.pi / 7,
                               lat1=np.pi / 2 - np.pi / 7, panels=32, order=16):
    """Φ(φ) - Φ(-π/2) for the geopotential in gradient balance with
    ``galewsky_zonal_wind``: -∫_{φ0}^{min(φ, φ1)} a u (2Ω sin s + u tan s / a) ds
    (the integrand is zero outside the jet), by ``panels`` Gauss-Legendre
    panels of ``order`` points each."""
    lat = np.asarray(lat, dtype=np.float64)
    x, w = np.polynomial.legendre.leggauss(order)
    top = np.clip(lat, lat0, lat1)
    # panel edges per latitude: lat0 + (top - lat0) * k / panels
    span = (top - lat0)[:, None] / panels                               # [nθ, 1]
    left = lat0 + span * np.arange(panels)[None, :]                     # [nθ, P]
    s = left[:, :, None] + 0.5 * span[:, :, None] * (x[None, None, :] + 1.0)   # [nθ, P, Q]
    u = galewsky_zonal_wind(s, u_max, lat0, lat1)
    f = radius * u * (2.0 * omega * np.sin(s) + u * np.tan(s) / radius)
    return -np.sum(f * w[None, None, :], axis=2).sum(axis=1) * 0.5 * span[:, 0]


def galewsky_geopotential(lat, wts, radius, omega, grav, mean_depth, u_max=80.0,
                          lat0=np.pi / 7, lat1=np.pi / 2 - np.pi / 7, panels=32, order=16):
    """Zonal geopotential on the latitudes ``lat`` (quadrature weights ``wts``):
    ``galewsky_geopotential_drop`` plus the constant that makes the
    Gaussian-quadrature mean Σ w Φ / Σ w equal g · mean_depth."""
    phi = galewsky_geopotential_drop(lat, radius, omega, u_max, lat0, lat1, panels, order)
    wts = np.asarray(wts, dtype=np.float64)
    return phi + grav * mean_depth - np.sum(wts * phi) / np.sum(wts)


def galewsky_bump(lat, lon, lon_c, bump=120.0, alpha=1.0 / 3.0, beta=1.0 / 15.0,
                  lat2=np.pi / 4):
    """Height perturbation ĥ cos φ exp(-(λ'/α)²) exp(-((φ2-φ)/β)²) in metres,
    λ' = λ - λ_c wrapped into [-π, π)."""
    lam = (np.asarray(lon)[None, :] - lon_c + np.pi) % (2 * np.pi) - np.pi
    lat = np.asarray(lat)[:, None]
    return bump * np.cos(lat) * np.exp(-(lam / alpha) ** 2) * np.exp(-((lat2 - lat) / beta) ** 2)


def Galewsky_Jet(mesh, dyn_data, grav=9.80616, omega=7.292e-5, mean_depth=1.0e4, seed=0,
                 bump_lon=None, perturb=True, **params):
    """Shallow-water state of the Galewsky jet on ``mesh`` (one level).

    Grid winds (u(φ), v = 0) go through Vor_Div_From_Grid_UV! to spectral
    vorticity and divergence, and back through UV_Grid_From_Vor_Div! /
    Trans_Spherical_To_Grid! to the truncated grid winds, vorticity and
    divergence; the geopotential Φ = gh (balanced zonal part plus g times the
    bump) goes through Trans_Grid_To_Spherical! to spe_lnps_c (also copied to
    spe_lnps_n, which Shallow_Water_Dynamics.jl:147 differentiates) and back
    to grid_ps_c; the previous-step fields are set equal to the current ones.
    The bump longitude is ``bump_lon`` or, when None, drawn
    uniformly from [0, 2π) with ``numpy.random.default_rng(seed)``.

    Returns ``{"h_0": g * mean_depth, "bump_lon": λ_c}``; h_0 is the
    reference geopotential of the semi-implicit correction.
    """
    p = dict(GALEWSKY)
    p.update(params)
    d = dyn_data
    if d.grid_u_c.shape[0] != 1:
        raise ValueError("Galewsky_Jet builds a one-level shallow-water state")
    lat, lon = np.asarray(mesh.θc), np.asarray(mesh.λc)
    radius = mesh.radius
    if bump_lon is None:
        bump_lon = float(np.random.default_rng(seed).uniform(0.0, 2 * np.pi))
    u = galewsky_zonal_wind(lat, p["u_max"], p["lat0"], p["lat1"])
    phi = galewsky_geopotential(lat, mesh.wts, radius, omega, grav, mean_depth, p["u_max"],
                                p["lat0"], p["lat1"])
    grid_phi = np.repeat(phi[:, None], lon.size, axis=1)
    if perturb:
        grid_phi = grid_phi + grav * galewsky_bump(lat, lon, bump_lon, p["bump"], p["alpha"],
                                                   p["beta"], p["lat2"])
    _host_to(d.grid_u_c, np.repeat(u[:, None], lon.size, axis=1)[None])
    d.grid_v_c.zero_()
    _host_to(d.grid_lnps, grid_phi[None])
    torch.cuda.synchronize()
    Vor_Div_From_Grid_UV(mesh, d.grid_u_c, d.grid_v_c, d.spe_vor_c, d.spe_div_c)
    Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    Trans_Spherical_To_Grid(mesh, d.spe_div_c, d.grid_div)
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    Trans_Grid_To_Spherical(mesh, d.grid_lnps, d.spe_lnps_c)
    Trans_Spherical_To_Grid(mesh, d.spe_lnps_c, d.grid_ps_c)
    mesh.synchronize()
    d.spe_lnps_n.copy_(d.spe_lnps_c)
    _previous_equals_current(d, ("vor", "div", "lnps"), ("u", "v", "ps"))
    torch.cuda.synchronize()
    return {"h_0": grav * mean_depth, "bump_lon": bump_lon}
