"""The HIP library against the committed fixtures (tests/golden/*.npz, the
oracle's outputs on seeded inputs): transforms at T21 and T42 over whole
arrays and at T85 through leading rows and per-field sums (1e-12 relative,
FP64; 1e-6 for the split-FP32 and BF16x3 Legendre stages), and short runs of
the three steppers (1e-11 barotropic and shallow water, 1e-10 Held-Suarez).
No oracle is imported: inputs are rebuilt from the seeds and formulas in
tests/golden/inputs.py."""
import os
import sys

import numpy as np
import pytest
import torch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

from inputs import (A, HS, OMEGA, T21, T42, T85, rh4_grid_vorticity, seeded_grid,  # noqa: E402
                    seeded_spec, sw_forcing)

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), device="cuda")


def load(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("name,cfg", [("golden_t21.npz", T21), ("golden_t42.npz", T42)])
def test_transforms_match_fixture(jgcm, name, cfg):
    z = load(name)
    N, nlat, nlev = cfg["N"], cfg["nlat"], cfg["nlev"]
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nlev, A)
    s = mesh.spectral_zeros()
    jgcm.Trans_Grid_To_Spherical(mesh, dev(z["grid_in"]), s)
    g = mesh.grid_zeros()
    jgcm.Trans_Spherical_To_Grid(mesh, dev(z["spec_in"]), g)
    torch.cuda.synchronize()
    assert rel(s.cpu().numpy(), z["spec_fwd"]) < 1e-12
    assert rel(g.cpu().numpy(), z["grid_inv"]) < 1e-12


def test_vor_div_and_winds_match_fixture(jgcm):
    z = load("golden_t21.npz")
    N, nlat = T21["N"], T21["nlat"]
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    vor, div = mesh.spectral_zeros(1), mesh.spectral_zeros(1)
    jgcm.Vor_Div_From_Grid_UV(mesh, dev(z["u_in"]), dev(z["v_in"]), vor, div)
    u, v = mesh.grid_zeros(1), mesh.grid_zeros(1)
    jgcm.UV_Grid_From_Vor_Div(mesh, dev(z["vor_in"]), dev(z["div_in"]), u, v)
    torch.cuda.synchronize()
    for got, key in ((vor, "vor_out"), (div, "div_out"), (u, "u_out"), (v, "v_out")):
        assert rel(got.cpu().numpy(), z[key]) < 1e-12, key


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("split32", 1e-6), ("bf16x3", 1e-6)])
def test_t85_transforms_match_fixture(jgcm, prec, tol):
    z = load("golden_t85.npz")
    N, nlat, nf = T85["N"], T85["nlat"], T85["nlev"]
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nf, A, precision=prec)
    s = mesh.spectral_zeros()
    jgcm.Trans_Grid_To_Spherical(mesh, dev(seeded_grid(nf, nlat, T85["seed"])), s)
    g = mesh.grid_zeros()
    jgcm.Trans_Spherical_To_Grid(mesh, dev(seeded_spec(N, nf, T85["seed"] + 1)), g)
    torch.cuda.synchronize()
    s, g = s.cpu().numpy(), g.cpu().numpy()
    assert rel(s[:, :6, :], z["spec_fwd_rows"]) < tol
    assert rel(g[:, list(T85["rows"]), :], z["grid_inv_rows"]) < tol
    assert rel((np.abs(s) ** 2).sum(axis=(1, 2)), z["spec_fwd_sumsq"]) < tol
    assert rel(s.sum(axis=(1, 2)), z["spec_fwd_sum"]) < tol
    assert rel((g ** 2).sum(axis=(1, 2)), z["grid_inv_sumsq"]) < tol
    assert rel(g.sum(axis=(1, 2)), z["grid_inv_sum"]) < tol * 10


def test_barotropic_run_matches_fixture(jgcm):
    z = load("golden_t21.npz")
    N, nlat = T21["N"], T21["nlat"]
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    d = jgcm.Dyn_Data("b", N, N + 1, 2 * nlat, nlat, 1, mesh=mesh)
    jgcm.Trans_Grid_To_Spherical(mesh, dev(rh4_grid_vorticity(mesh.sinθ, mesh.cosθ, mesh.λc)),
                                 d.spe_vor_c)
    jgcm.Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    jgcm.UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    for i in range(T21["baro_steps"]):
        jgcm.Barotropic_Dynamics(mesh, atmo, d, I)
        if i == 0:
            jgcm.Update_Init_Step(I)
    torch.cuda.synchronize()
    assert rel(d.spe_vor_c.cpu().numpy(), z["baro_vor_c"]) < 1e-11
    assert rel(d.spe_vor_p.cpu().numpy(), z["baro_vor_p"]) < 1e-11
    assert rel(d.grid_u_c.cpu().numpy(), z["baro_u_c"]) < 1e-11


def test_shallow_water_run_matches_fixture(jgcm):
    z = load("golden_t21.npz")
    N, nlat = T21["N"], T21["nlat"]
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    d = jgcm.Dyn_Data("sw", N, N + 1, 2 * nlat, nlat, 1, mesh=mesh)
    h_0, h_eq = sw_forcing(mesh.θc, mesh.λc)
    d.grid_lnps.fill_(h_0)
    jgcm.Trans_Grid_To_Spherical(mesh, d.grid_lnps, d.spe_lnps_c)
    atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, mesh.laplacian_eig, 0.5, 1200, True, 0, 691200)
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
    heq = dev(h_eq)
    for i in range(T21["sw_steps"]):
        jgcm.Shallow_Water_Physics(d, km, kt, heq)
        jgcm.Shallow_Water_Dynamics(mesh, atmo, h_0, d, I)
        if i == 0:
            jgcm.Update_Init_Step(I)
    torch.cuda.synchronize()
    assert rel(d.spe_vor_c.cpu().numpy(), z["sw_vor_c"]) < 1e-11
    assert rel(d.spe_div_c.cpu().numpy(), z["sw_div_c"]) < 1e-11
    assert rel(d.spe_lnps_c.cpu().numpy(), z["sw_h_c"]) < 1e-11


def test_held_suarez_run_matches_fixture(jgcm):
    z = load("golden_t21.npz")
    N, nlat, nd, dt = T21["N"], T21["nlat"], T21["hs_nd"], T21["hs_dt"]
    nlon = 2 * nlat
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, nlon, nd, A)
    vc = jgcm.Vert_Coordinate(nlon, nlat, nd, "even_sigma", "simmons_and_burridge",
                              "second_centered_wts", 1e5, mesh=mesh)
    atmo = jgcm.Atmo_Data("hs", nlon, nlat, nd, True, True, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.15741e-4, mesh.laplacian_eig, 0.5, dt, True, 0, 86400 * 200)
    si = jgcm.Semi_Implicit_Solver(vc, atmo, I, 1e5, np.full(nd, 300.0), mesh.wave_numbers)
    d = jgcm.Dyn_Data("hs", N, N + 1, nlon, nlat, nd, mesh=mesh)
    jgcm.Get_Topography(d.grid_geopots)
    noise = 0.1 * np.random.default_rng(T21["hs_seed"]).standard_normal((nd, nlat, nlon))
    jgcm.Spectral_Initialize_Fields(mesh, atmo, vc, 1.0e5, 300.0, d.grid_geopots, d,
                                    grid_t_noise=dev(noise))
    for i in range(T21["hs_steps"]):
        jgcm.Atmosphere_Update(mesh, atmo, vc, si, d, HS)
        if i == 0:
            jgcm.Update_Init_Step(si)
    torch.cuda.synchronize()
    for f, key in (("spe_vor_c", "hs_vor_c"), ("spe_div_c", "hs_div_c"), ("spe_t_c", "hs_t_c"),
                   ("spe_lnps_c", "hs_lnps_c")):
        assert rel(getattr(d, f).cpu().numpy(), z[key]) < 1e-10, f
