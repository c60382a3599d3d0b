"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no golden vectors (it has no test suite) and cannot run
here, so the fixtures are the oracle's outputs on seeded inputs; the oracle
itself is pinned by closed forms (tests/test_oracle_*.py).  They freeze the
oracle's behaviour (test_golden.py re-derives them on the CPU) and give the
GPU tests fixed expected values (test_gpu_golden.py).

    python tests/golden/make_golden.py      # from the repository root

Contents (all FP64; spectral arrays in the reference layout [nlev][N+2][N+1]):
  golden_t21.npz  transforms (full arrays, inputs included), vor/div <-> u/v,
                  6 barotropic steps (RH4), 6 shallow-water steps
                  (exp/Shallow_Water forcing), 3 Held-Suarez steps (L10)
  golden_t42.npz  transforms, one level-field each way (full arrays)
  golden_t85.npz  transforms of 4 seeded level-fields: leading spectral rows,
                  grid rows at both poles and the equator, and per-field sums
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import FilteredLeapfrog, SpectralSphericalMesh  # noqa: E402
from oracle import dynamics as odyn  # noqa: E402
from oracle import primitive as P  # noqa: E402

from inputs import (A, HS, OMEGA, T21, T42, T85, rh4_grid_vorticity, seeded_grid,  # noqa: E402
                    seeded_spec, sw_forcing)


def transforms(N, nlat, nlev, seed):
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, nlev, A)
    g = seeded_grid(nlev, nlat, seed)
    s = seeded_spec(N, nlev, seed + 1)
    fwd = M.trans_grid_to_spherical(g, M.spec_zeros())
    inv = M.trans_spherical_to_grid(s, M.grid_zeros())
    return M, g, s, fwd, inv


def rh4(M, N):
    s = np.zeros((1, N + 2, N + 1), dtype=complex)
    M.trans_grid_to_spherical(rh4_grid_vorticity(M.sin_theta, M.cos_theta, M.lam_c), s)
    return s


def barotropic_run(N, nlat, steps):
    """exp/Barotropic: RH4 vorticity, Δt = 1800 s, 0.1-day ∇⁸ damping."""
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    d = odyn.DynData("b", N, N + 1, 2 * nlat, nlat, 1)
    d.spe_vor_c[...] = rh4(M, N)
    M.trans_spherical_to_grid(d.spe_vor_c, d.grid_vor)
    M.uv_grid_from_vor_div(d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)
    atmo = odyn.AtmoData("b", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    I = FilteredLeapfrog(0.04, 4, 1.0 / (0.1 * 86400), M.laplacian_eig, 0.0, 1800, True, 0, 864000)
    for i in range(steps):
        odyn.barotropic_dynamics(M, atmo, d, I)
        if i == 0:
            I.update_init_step()
    return d


def sw_run(N, nlat, steps):
    """exp/Shallow_Water (SW.jl:57-95): u = v = 0, h = h_0, forcing toward h_eq."""
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    d = odyn.DynData("sw", N, N + 1, 2 * nlat, nlat, 1)
    h_0, h_eq = sw_forcing(M.theta_c, M.lam_c)
    d.grid_lnps[...] = h_0
    M.trans_grid_to_spherical(d.grid_lnps, d.spe_lnps_c)
    atmo = odyn.AtmoData("sw", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    I = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.5, 1200, True, 0, 691200)
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
    for i in range(steps):
        odyn.shallow_water_physics(d, km, kt, h_eq)
        odyn.shallow_water_dynamics(M, atmo, h_0, d, I)
        if i == 0:
            I.update_init_step()
    return d


def hs_run(N, nlat, nd, dt, steps, seed=11):
    """exp/HSt42 step sequence (Atmosphere_Update!) from the seeded-noise initial state."""
    nlon = 2 * nlat
    M = SpectralSphericalMesh(N, N + 1, nlat, nlon, nd, A)
    vc = P.VertCoordinate(nlon, nlat, nd, "even_sigma", "simmons_and_burridge", "second_centered_wts", 1e5)
    atmo = odyn.AtmoData("hs", nlon, nlat, nd, True, True, False, False, M.sin_theta, A, OMEGA)
    I = FilteredLeapfrog(0.04, 4, 1.15741e-4, M.laplacian_eig, 0.5, dt, True, 0, 86400 * 200)
    si = P.SemiImplicitSolver(vc, atmo, I, 1e5, np.full(nd, 300.0), M.wave_numbers)
    d = odyn.DynData("hs", N, N + 1, nlon, nlat, nd)
    noise = 0.1 * np.random.default_rng(seed).standard_normal((nd, nlat, nlon))
    P.spectral_initialize_fields(M, atmo, vc, 1.0e5, 300.0, d.grid_geopots, d, grid_t_noise=noise)
    for i in range(steps):
        P.atmosphere_update(M, atmo, vc, si, d, HS)
        if i == 0:
            si.update_init_step()
    return d


def make_t21():
    c = T21
    N, nlat = c["N"], c["nlat"]
    M, g, s, fwd, inv = transforms(N, nlat, c["nlev"], c["seed"])
    u, v = (seeded_grid(1, nlat, k) for k in c["uv_seeds"])
    vor, div = M.vor_div_from_grid_uv(u, v, np.zeros((1, N + 2, N + 1), complex),
                                      np.zeros((1, N + 2, N + 1), complex))
    sv, sd = (seeded_spec(N, 1, k) for k in c["vd_seeds"])
    gu, gv = np.zeros((1, nlat, 2 * nlat)), np.zeros((1, nlat, 2 * nlat))
    M.uv_grid_from_vor_div(sv, sd, gu, gv)
    b = barotropic_run(N, nlat, c["baro_steps"])
    w = sw_run(N, nlat, c["sw_steps"])
    h = hs_run(N, nlat, c["hs_nd"], c["hs_dt"], c["hs_steps"], c["hs_seed"])
    return dict(grid_in=g, spec_fwd=fwd, spec_in=s, grid_inv=inv, u_in=u, v_in=v, vor_out=vor,
                div_out=div, vor_in=sv, div_in=sd, u_out=gu, v_out=gv,
                baro_vor_c=b.spe_vor_c, baro_vor_p=b.spe_vor_p, baro_u_c=b.grid_u_c,
                sw_vor_c=w.spe_vor_c, sw_div_c=w.spe_div_c, sw_h_c=w.spe_lnps_c,
                hs_vor_c=h.spe_vor_c, hs_div_c=h.spe_div_c, hs_t_c=h.spe_t_c, hs_lnps_c=h.spe_lnps_c)


def make_t42():
    c = T42
    M, g, s, fwd, inv = transforms(c["N"], c["nlat"], c["nlev"], c["seed"])
    return dict(grid_in=g, spec_fwd=fwd, spec_in=s, grid_inv=inv)


def make_t85():
    c = T85
    M, g, s, fwd, inv = transforms(c["N"], c["nlat"], c["nlev"], c["seed"])
    rows = list(c["rows"])
    return dict(spec_fwd_rows=fwd[:, :6, :], spec_fwd_sumsq=(np.abs(fwd) ** 2).sum(axis=(1, 2)),
                spec_fwd_sum=fwd.sum(axis=(1, 2)), grid_inv_rows=inv[:, rows, :],
                grid_inv_sumsq=(inv ** 2).sum(axis=(1, 2)), grid_inv_sum=inv.sum(axis=(1, 2)))


BUILDERS = {"golden_t21.npz": make_t21, "golden_t42.npz": make_t42, "golden_t85.npz": make_t85}

if __name__ == "__main__":
    for name, fn in BUILDERS.items():
        np.savez_compressed(os.path.join(HERE, name), **fn())
        print("wrote", name, os.path.getsize(os.path.join(HERE, name)), "bytes")
