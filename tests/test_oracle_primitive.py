"""Pins for oracle.primitive (CPU): the Held-Suarez primitive-equation step.

Reference: Press_And_Geopot.jl, Vert_Coordinate.jl, HS_Forcing.jl,
Semi_Implicit.jl, Spectral_Dynamics.jl.  Independent checks: closed forms
(isothermal hydrostatic column, Simmons-Burridge ln p, Held-Suarez 1994
T_eq / k_T / k_v), the semi-implicit linear operators against the nonlinear
reference functions they linearise (Four_In_One!, Compute_Geopotential!,
Pressure_Variables!), W_n (I - ξ² eig D) = I, a resting isothermal atmosphere
staying at rest, and mass conservation by the corrections.
"""
import json
import os

import numpy as np
import pytest

from oracle import FilteredLeapfrog, SpectralSphericalMesh
from oracle import dynamics as odyn
from oracle import primitive as P

R, KAPPA = 287.04, 2.0 / 7.0
HS = {"σ_b": 0.7, "k_f": 1.0, "k_a": 1.0 / 40.0, "k_s": 1.0 / 4.0, "ΔT_y": 60.0, "Δθ_z": 10.0}
# the coordinates with their level counts (even sigma, uneven sigma and hybrid: any nd)
COORDS = [("even_sigma", 10), ("mcm", 14), ("v197", 18), ("uneven_sigma", 20), ("hybrid", 20)]
SIGMA_TABLES = os.path.join(os.path.dirname(__file__), "golden", "sigma_tables.json")


def column(nd=10, ps=1.0e5, coord="even_sigma"):
    vc = P.VertCoordinate(1, 1, nd, coord, "simmons_and_burridge", "second_centered_wts", 1.0e5)
    sh = lambda k: np.zeros((k, 1, 1))
    ph, Δp, lh, pf, lf = sh(nd + 1), sh(nd), sh(nd + 1), sh(nd), sh(nd)
    P.pressure_variables(vc, np.full((1, 1, 1), ps), ph, Δp, lh, pf, lf)
    return vc, ph, Δp, lh, pf, lf


def test_pressure_variables_closed_forms():
    nd, ps = 10, 98765.0
    vc, ph, Δp, lh, pf, lf = column(nd, ps)
    sig = np.linspace(0, 1, nd + 1)
    np.testing.assert_allclose(ph[:, 0, 0], sig * ps, rtol=1e-15)
    # Simmons-Burridge: ln p_k = (p+ ln p+ - p- ln p-)/Δp - 1, with 0 ln 0 = 0 at the top
    pp, pm = ph[1:, 0, 0], ph[:-1, 0, 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        pm_ln = np.where(pm > 0, pm * np.log(np.where(pm > 0, pm, 1)), 0.0)
    expect = (pp * np.log(pp) - pm_ln) / (pp - pm) - 1.0
    np.testing.assert_allclose(lf[:, 0, 0], expect, rtol=1e-14)
    np.testing.assert_allclose(pf[:, 0, 0], np.exp(expect), rtol=1e-14)
    assert lh[0, 0, 0] == 0.0


def test_sigma_tables_match_the_reference_text():
    """The mcm and v197 tables equal the fixture read from the reference's
    source text; v197 is symmetric about σ = 1/2 to its 7 printed digits;
    table coordinates with the wrong level count and unknown options raise."""
    with open(SIGMA_TABLES) as f:
        tables = json.load(f)
    for name, t in tables.items():
        vc = P.VertCoordinate(1, 1, t["nd"], name, "simmons_and_burridge", "second_centered_wts")
        assert list(vc.bk) == t["b"] and not vc.ak.any()
        assert np.all(np.diff(vc.bk) > 0) and vc.bk[0] == 0.0 and vc.bk[-1] == 1.0
        with pytest.raises(ValueError):
            P.VertCoordinate(1, 1, t["nd"] - 1, name, "simmons_and_burridge", "second_centered_wts")
    b = np.array(tables["v197"]["b"])
    np.testing.assert_allclose(b + b[::-1], 1.0, atol=1.5e-7)
    with pytest.raises(ValueError):
        P.VertCoordinate(1, 1, 10, "sigma_and_more", "simmons_and_burridge", "second_centered_wts")


@pytest.mark.parametrize("surf_res,exponent", [(1.0, 2.5), (0.2, 2.5), (0.5, 3.0)])
def test_uneven_sigma_closed_form(surf_res, exponent):
    """Compute_Uneven_Sigma (Vert_Coordinate.jl:120-141, the docstring's form):
    with surf_res = 1 the levels are even in ln σ, ln b_k = -H (1 - k/nd);
    otherwise ln b_k = -H (s ζ + (1-s) ζ^e) at ζ = 1 - k/nd, checked at a few
    levels by hand; b rises monotonically from 0 (zero top) to 1; a = 0."""
    nd, H = 20, 4.0
    vc = P.VertCoordinate(1, 1, nd, "uneven_sigma", "simmons_and_burridge", "second_centered_wts",
                          surf_res=surf_res, exponent=exponent)
    b = vc.bk
    assert b[0] == 0.0 and b[nd] == 1.0 and not vc.ak.any()
    assert np.all(np.diff(b) > 0)
    for k in (1, 5, 10, 19):
        zeta = 1.0 - k / nd
        expect = -H * (surf_res * zeta + (1.0 - surf_res) * zeta ** exponent)
        assert abs(np.log(b[k]) - expect) < 1e-13, k
    if surf_res == 1.0:
        np.testing.assert_allclose(np.diff(np.log(b[1:])), H / nd, rtol=1e-12)
    top = P.VertCoordinate(1, 1, nd, "uneven_sigma", "simmons_and_burridge", "second_centered_wts",
                           zero_top=False)
    assert abs(top.bk[0] - np.exp(-H)) < 1e-16


def test_hybrid_coordinate_blends_pressure_and_sigma_levels():
    """hybrid (Vert_Coordinate.jl:71-77 with Transition :93-110): at the
    reference surface pressure every half level sits where the uneven σ level
    does (a_k + b_k p_ref = p_ref b_u,k); above σ = p_press the levels are pure
    pressure (b = 0), below σ = p_sigma pure σ (a = 0), and between, the σ part
    is sin² of the distance across the band."""
    nd, p_ref = 20, 101325.0
    vc = P.VertCoordinate(1, 1, nd, "hybrid", "simmons_and_burridge", "second_centered_wts",
                          p_ref=p_ref)
    _, bu = P.uneven_sigma(nd, 4.0, 1.0, 2.5, False)
    np.testing.assert_allclose(vc.ak + vc.bk * p_ref, p_ref * bu, rtol=1e-15)
    top, low = bu <= 0.1, bu >= 0.3
    assert top.any() and low.any() and (~top & ~low).any()
    assert not vc.bk[top].any() and not vc.ak[low].any()
    np.testing.assert_allclose(vc.ak[top], p_ref * bu[top], rtol=1e-15)
    mid = ~top & ~low
    np.testing.assert_allclose(vc.bk[mid] / bu[mid],
                               np.sin(0.5 * np.pi * (bu[mid] - 0.1) / 0.2) ** 2, rtol=1e-14)
    assert vc.bk[-1] == 1.0 and vc.ak[-1] == 0.0
    # half-level pressures increase downward over the range of surface
    # pressures a Held-Suarez run visits (at ps below ≈0.57 p_ref the blend
    # would fold levels over)
    for ps in (6.0e4, 1.0e5, 1.1e5):
        assert np.all(np.diff(vc.ak + vc.bk * ps) > 0)


@pytest.mark.parametrize("coord,nd", [("even_sigma", 12), ("mcm", 14), ("v197", 18)])
def test_isothermal_geopotential_exact(coord, nd):
    T0 = 263.0
    vc, ph, Δp, lh, pf, lf = column(nd, coord=coord)
    atmo = odyn.AtmoData("hs", 1, 1, nd, False, False, False, False, np.zeros(1), 6.371e6)
    gh, gf = np.zeros((nd + 1, 1, 1)), np.zeros((nd, 1, 1))
    P.compute_geopotential(vc, atmo, lh, lf, np.full((nd, 1, 1), T0), np.zeros((1, 1, 1)), gf, gh)
    ps = ph[nd, 0, 0]
    np.testing.assert_allclose(gh[1:, 0, 0], R * T0 * np.log(ps / ph[1:, 0, 0]), rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(gf[:, 0, 0], R * T0 * (np.log(ps) - lf[:, 0, 0]), rtol=1e-13)


def test_held_suarez_closed_forms():
    nd = 10
    sin = np.array([-0.7, 0.1, 0.5])
    nlat = sin.size
    vc = P.VertCoordinate(2, nlat, nd, "even_sigma", "simmons_and_burridge", "second_centered_wts", 1e5)
    ps = np.full((1, nlat, 2), 1.0e5)
    ph, Δp, lh = np.zeros((nd + 1, nlat, 2)), np.zeros((nd, nlat, 2)), np.zeros((nd + 1, nlat, 2))
    pf, lf = np.zeros((nd, nlat, 2)), np.zeros((nd, nlat, 2))
    P.pressure_variables(vc, ps, ph, Δp, lh, pf, lf)
    atmo = odyn.AtmoData("hs", 2, nlat, nd, False, False, False, False, sin, 6.371e6)
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal((2, nd, nlat, 2)) * 10
    t = 250 + rng.standard_normal((nd, nlat, 2)) * 10
    δu, δv, δt, teq = (np.zeros((nd, nlat, 2)) for _ in range(4))
    P.hs_forcing(atmo, 600, sin, u, v, ph, pf, t, δu, δv, teq, δt, HS)
    σ = pf / 1e5
    s2 = (sin ** 2)[None, :, None]
    teq_x = np.maximum(200.0, (315 - 60 * s2 - 10 * (1 - s2) * np.log(σ)) * σ ** KAPPA)
    np.testing.assert_allclose(teq, teq_x, rtol=1e-14)
    kv = (1 / 86400) * np.maximum(0, (σ - 0.7) / 0.3)
    np.testing.assert_allclose(δu, -kv * u, rtol=1e-13, atol=1e-20)
    kt = 1 / (40 * 86400) + (1 / (4 * 86400) - 1 / (40 * 86400)) * np.maximum(0, (σ - 0.7) / 0.3) * (1 - s2) ** 2
    heat = -((u + 0.5 * δu * 600) * δu + (v + 0.5 * δv * 600) * δv) / (R / KAPPA)
    np.testing.assert_allclose(δt, heat - kt * (t - teq_x), rtol=1e-12, atol=1e-18)


def _solver(nd=10, N=21, xi_dt=600, coord="even_sigma"):
    M = SpectralSphericalMesh(N, N + 1, 32, 64, nd, 6.371e6)
    vc = P.VertCoordinate(64, 32, nd, coord, "simmons_and_burridge", "second_centered_wts", 1e5)
    atmo = odyn.AtmoData("hs", 64, 32, nd, True, True, False, False, M.sin_theta, 6.371e6)
    I = FilteredLeapfrog(0.04, 4, 1.15741e-4, M.laplacian_eig, 0.5, xi_dt, True, 0, 86400)
    si = P.SemiImplicitSolver(vc, atmo, I, 1e5, np.full(nd, 300.0), M.wave_numbers)
    return M, vc, atmo, I, si


@pytest.mark.parametrize("coord,nd", COORDS)
def test_semi_implicit_linear_operators_match_nonlinear(coord, nd):
    M, vc, atmo, I, si = _solver(nd, coord=coord)
    shp = (1, 1)
    T0 = 300.0
    # resting isothermal column at ps_ref: Four_In_One's δps, δT response to a
    # unit divergence at level k is -nu[k] and -tau[:, k] (vertical advection
    # of an isothermal profile vanishes)
    _, ph, Δp, lh, pf, lf = column(nd, 1e5, coord)
    for k in range(nd):
        div = np.zeros((nd, 1, 1))
        div[k] = 1.0
        z = lambda n: np.zeros((n, 1, 1))
        Mh, w, δu, δv, δps, δt = z(nd + 1), z(nd), z(nd), z(nd), z(1), z(nd)
        P.four_in_one(vc, atmo, div, z(nd), z(nd), np.full((1, 1, 1), 1e5), Δp, lh, lf, pf, z(1),
                      z(1), np.full((nd, 1, 1), T0), Mh, w, δu, δv, δps, δt)
        assert δps[0, 0, 0] == pytest.approx(-si.nu[k], rel=1e-13)
        # entries near zero come from cancellation: absolute bound on the column's scale
        np.testing.assert_allclose(δt[:, 0, 0], -si.tau[:, k], rtol=1e-12,
                                   atol=1e-13 * np.abs(si.tau[:, k]).max())
    # gamma: Φ is linear in T at fixed pressure
    gh, g0, g1 = np.zeros((nd + 1, 1, 1)), np.zeros((nd, 1, 1)), np.zeros((nd, 1, 1))
    dT = np.random.default_rng(1).standard_normal((nd, 1, 1))
    P.compute_geopotential(vc, atmo, lh, lf, np.full((nd, 1, 1), T0), np.zeros((1, 1, 1)), g0, gh)
    P.compute_geopotential(vc, atmo, lh, lf, T0 + dT, np.zeros((1, 1, 1)), g1, gh)
    dg = (g1 - g0)[:, 0, 0]   # a difference of O(1e5) values: absolute bound on its scale
    np.testing.assert_allclose(si.gamma @ dT[:, 0, 0], dg, rtol=1e-10, atol=1e-12 * np.abs(dg).max())
    # h = R T dlnp_full/dps + dΦ/dps by central differences in ps
    eps = 1.0
    vals = []
    for s in (1e5 + eps, 1e5 - eps):
        _, ph2, Δp2, lh2, pf2, lf2 = column(nd, s, coord)
        g = np.zeros((nd, 1, 1))
        P.compute_geopotential(vc, atmo, lh2, lf2, np.full((nd, 1, 1), T0), np.zeros((1, 1, 1)), g,
                               np.zeros((nd + 1, 1, 1)))
        vals.append((g[:, 0, 0] + R * T0 * lf2[:, 0, 0]))
    fd = (vals[0] - vals[1]) / (2 * eps)
    np.testing.assert_allclose(si.h, fd, rtol=1e-6)


@pytest.mark.parametrize("coord,nd", COORDS)
def test_wave_matrices_invert_and_gravity_speeds_positive(coord, nd):
    M, vc, atmo, I, si = _solver(nd, coord=coord)
    xi = I.get_xi()
    for n in range(si.num_wavenumbers + 1):
        A = np.eye(si.nd) - xi ** 2 * M.laplacian_eig[0, n] * si.div_mat
        np.testing.assert_allclose(si.wave_matrix[n] @ A, np.eye(si.nd), atol=1e-12)
    ev = np.linalg.eigvals(si.div_mat)
    assert np.all(ev.real > 0) and np.abs(ev.imag).max() < 1e-6 * ev.real.max()
    # equivalent depth of the fastest mode ~ R T / g -> c ~ 300 m/s
    assert 2e4 < ev.real.max() < 2e5


def _hs_state(N=21, nlat=32, nd=10, noise=0.1, seed=3, coord="even_sigma"):
    M, vc, atmo, I, si = _solver(nd, N, coord=coord)
    d = odyn.DynData("hs", N, N + 1, 2 * nlat, nlat, nd)
    rng = np.random.default_rng(seed)
    tn = noise * rng.standard_normal((nd, nlat, 2 * nlat)) if noise else None
    P.spectral_initialize_fields(M, atmo, vc, 1.0e5, 300.0, d.grid_geopots, d, grid_t_noise=tn)
    return M, vc, atmo, I, si, d


@pytest.mark.parametrize("coord,nd", COORDS)
def test_resting_isothermal_atmosphere_stays_at_rest(coord, nd):
    M, vc, atmo, I, si, d = _hs_state(nd=nd, noise=0.0, coord=coord)
    d.spe_vor_c[...] = 0
    d.spe_vor_p[...] = 0
    d.grid_u_c[...] = d.grid_v_c[...] = d.grid_u_p[...] = d.grid_v_p[...] = 0
    d.grid_vor[...] = 0
    off = {k: 0.0 for k in HS}
    off["σ_b"] = 0.7
    for i in range(3):
        P.atmosphere_update(M, atmo, vc, si, d, off)
        if i == 0:
            si.update_init_step()
    # round-off only (the flow speeds of the model are O(1-100) m/s)
    assert np.abs(d.grid_u_c).max() < 1e-8 and np.abs(d.grid_v_c).max() < 1e-8
    np.testing.assert_allclose(d.grid_t_c, 300.0, rtol=1e-11)
    np.testing.assert_allclose(d.grid_ps_c, 1e5, rtol=1e-11)


@pytest.mark.parametrize("coord,nd", COORDS)
def test_hs_steps_conserve_mass(coord, nd):
    M, vc, atmo, I, si, d = _hs_state(nd=nd, coord=coord)
    m0 = M.area_weighted_global_mean(d.grid_ps_c)
    for i in range(4):
        P.atmosphere_update(M, atmo, vc, si, d, HS)
        if i == 0:
            si.update_init_step()
    assert M.area_weighted_global_mean(d.grid_ps_c) == pytest.approx(m0, rel=1e-13)
    assert np.all(np.isfinite(d.grid_t_c)) and np.abs(d.grid_u_c).max() < 5.0
