"""Barotropic and shallow-water steppers and their state (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:

* ``AtmoData``            <- ``Atmo_Data`` ctor, Atmo_Data.jl:34-55 (g defaults to 9.80)
* ``compute_abs_vor``     <- ``Compute_Abs_Vor!`` Atmo_Data.jl:59-65
* ``DynData``             <- ``Dyn_Data`` ctor, Dyn_Data.jl:158-324 (prognostic and
                             scratch fields used by the steppers)
* ``time_advance``        <- ``Time_Advance!`` Dyn_Data.jl:326-364 (``_n`` untouched)
* ``barotropic_dynamics`` <- ``Barotropic_Dynamics!`` Barotropic_Dynamics.jl:2-45
* ``shallow_water_physics`` <- ``Shallow_Water_Physics!`` Shallow_Water_Dynamics.jl:7-21
* ``sw_implicit_correction`` <- SW ``Implicit_Correction!`` Shallow_Water_Dynamics.jl:25-54
* ``shallow_water_dynamics`` <- ``Shallow_Water_Dynamics!`` Shallow_Water_Dynamics.jl:88-179

Pins (tests/test_oracle_dynamics.py): the Rossby-Haurwitz wave-4 solution of
the nondivergent vorticity equation (exact travelling solution), a zonal
flow in solid-body rotation staying steady in the shallow-water model (the
geostrophically balanced state), the 2x2 linear system of the SW implicit
correction solved directly.
"""
import numpy as np


class AtmoData:
    def __init__(self, name, nlon, nlat, nd, do_mass_correction, do_energy_correction,
                 do_water_correction, use_virtual_temperature, sin_theta, radius,
                 omega=7.292e-5, grav=9.80, rdgas=287.04, kappa=2.0 / 7.0, rvgas=461.50):
        self.name = name
        self.nlon, self.nlat, self.nd = nlon, nlat, nd
        self.do_mass_correction = do_mass_correction
        self.do_energy_correction = do_energy_correction
        self.do_water_correction = do_water_correction
        self.use_virtual_temperature = use_virtual_temperature
        self.radius, self.omega, self.grav = radius, omega, grav
        self.rdgas, self.kappa, self.rvgas = rdgas, kappa, rvgas
        self.coriolis = 2 * omega * np.asarray(sin_theta)
        self.cp_air = rdgas / kappa


def compute_abs_vor(grid_vor, coriolis, grid_absvor):
    grid_absvor[...] = grid_vor + coriolis[None, :, None]
    return grid_absvor


class DynData:
    """The Dyn_Data fields the steppers touch, zero-initialised (Dyn_Data.jl:158-324)."""

    SPEC3 = ["vor", "div", "t"]
    SPEC1 = ["lnps"]

    def __init__(self, name, num_fourier, num_spherical, nlon, nlat, nd, num_grid_tracters=0,
                 num_spe_tracters=0):
        self.name = name
        self.nd = nd
        self.num_grid_tracters, self.num_spe_tracters = num_grid_tracters, num_spe_tracters
        sp = lambda k: np.zeros((k, num_spherical + 1, num_fourier + 1), np.complex128)
        gr = lambda k: np.zeros((k, nlat, nlon))
        # tracers (Dyn_Data.jl:181-184, 208-211): [tracer][level][...]
        for s in "ncp":
            setattr(self, f"spe_tracers_{s}", np.zeros(
                (num_spe_tracters, nd, num_spherical + 1, num_fourier + 1), np.complex128))
            setattr(self, f"grid_tracers_{s}", np.zeros((num_grid_tracters, nd, nlat, nlon)))
        for f in self.SPEC3:
            for s in "ncp":
                setattr(self, f"spe_{f}_{s}", sp(nd))
        for s in "ncp":
            setattr(self, f"spe_lnps_{s}", sp(1))
        for f in ("u", "v", "t"):
            for s in "ncp":
                setattr(self, f"grid_{f}_{s}", gr(nd))
        for s in "ncp":
            setattr(self, f"grid_ps_{s}", gr(1))
        self.spe_δvor, self.spe_δdiv, self.spe_δt = sp(nd), sp(nd), sp(nd)
        self.spe_δlnps = sp(1)
        self.grid_vor, self.grid_div = gr(nd), gr(nd)
        self.grid_δu, self.grid_δv, self.grid_δt = gr(nd), gr(nd), gr(nd)
        self.grid_lnps, self.grid_δlnps, self.grid_δps = gr(1), gr(1), gr(1)
        self.grid_p_full, self.grid_lnp_full, self.grid_Δp = gr(nd), gr(nd), gr(nd)
        self.grid_p_half, self.grid_lnp_half = gr(nd + 1), gr(nd + 1)
        self.grid_dλ_ps, self.grid_dθ_ps = gr(1), gr(1)
        self.grid_absvor = gr(nd)
        self.grid_w_full, self.grid_M_half = gr(nd), gr(nd + 1)
        self.spe_energy, self.grid_energy_full = sp(nd), gr(nd)
        self.grid_geopots, self.grid_geopot_full, self.grid_geopot_half = gr(1), gr(nd), gr(nd + 1)
        self.grid_z_full, self.grid_z_half = gr(nd), gr(nd + 1)
        self.grid_t_eq = gr(nd)
        self.spe_d1, self.spe_d2 = sp(nd), sp(nd)
        self.grid_d_full1, self.grid_d_full2 = gr(nd), gr(nd)
        self.spe_zeros = sp(nd)


def time_advance(d):
    """Dyn_Data.jl:326-364: p <- c, c <- n for the prognostic fields; n untouched."""
    for f in ("vor", "div", "lnps", "t"):
        getattr(d, f"spe_{f}_p")[...] = getattr(d, f"spe_{f}_c")
        getattr(d, f"spe_{f}_c")[...] = getattr(d, f"spe_{f}_n")
    for f in ("u", "v", "ps", "t"):
        getattr(d, f"grid_{f}_p")[...] = getattr(d, f"grid_{f}_c")
        getattr(d, f"grid_{f}_c")[...] = getattr(d, f"grid_{f}_n")
    # tracers only when there are any (:341-344, :360-363)
    for f in ("spe", "grid"):
        if getattr(d, f"num_{f}_tracters", 0) > 0:
            getattr(d, f"{f}_tracers_p")[...] = getattr(d, f"{f}_tracers_c")
            getattr(d, f"{f}_tracers_c")[...] = getattr(d, f"{f}_tracers_n")


def barotropic_dynamics(mesh, atmo, d, integ):
    """Barotropic_Dynamics.jl:2-45 (the @show norm line is diagnostics only)."""
    compute_abs_vor(d.grid_vor, atmo.coriolis, d.grid_absvor)
    d.grid_δu[...] = d.grid_absvor * d.grid_v_c
    d.grid_δv[...] = -d.grid_absvor * d.grid_u_c
    mesh.vor_div_from_grid_uv(d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    integ.compute_spectral_damping(d.spe_vor_c, d.spe_vor_p, d.spe_δvor)
    integ.filtered_leapfrog(d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)
    mesh.trans_spherical_to_grid(d.spe_vor_n, d.grid_vor)
    mesh.uv_grid_from_vor_div(d.spe_vor_n, d.spe_zeros, d.grid_u_n, d.grid_v_n)
    time_advance(d)


def shallow_water_physics(d, kappa_m, kappa_t, h_eq):
    """Shallow_Water_Dynamics.jl:7-21 (h lives in grid_ps_c, δh in grid_δlnps)."""
    d.grid_δu[...] = -kappa_m * d.grid_u_c
    d.grid_δv[...] = -kappa_m * d.grid_v_c
    d.grid_δlnps[...] = -kappa_t * (d.grid_ps_c - h_eq)


def sw_implicit_correction(integ, div_c, div_p, h_c, h_p, h_0, δdiv, δh):
    """Shallow_Water_Dynamics.jl:25-54."""
    eig = integ.laplacian_eigen.T[None, :, :]
    dt = integ.dt if integ.init_step else 2.0 * integ.dt
    δdiv += eig * (h_c - h_p)
    δh += h_0 * (div_c - div_p)
    mu = integ.implicit_coef * dt
    mu2 = mu ** 2
    δdiv[...] = (δdiv - mu * eig * δh) / (1.0 - mu2 * eig * h_0)
    δh -= mu * h_0 * δdiv


def shallow_water_dynamics(mesh, atmo, h_0, d, integ):
    """Shallow_Water_Dynamics.jl:88-179."""
    compute_abs_vor(d.grid_vor, atmo.coriolis, d.grid_absvor)
    d.grid_δu += d.grid_absvor * d.grid_v_c
    d.grid_δv += -d.grid_absvor * d.grid_u_c
    mesh.vor_div_from_grid_uv(d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    grid_kin, spe_kin = d.grid_d_full1, d.spe_d1
    grid_kin[...] = 0.5 * (d.grid_u_c ** 2 + d.grid_v_c ** 2)
    mesh.trans_grid_to_spherical(grid_kin, spe_kin)
    d.spe_energy[...] = spe_kin + d.spe_lnps_c
    mesh.apply_laplacian(d.spe_energy)
    d.spe_δdiv -= d.spe_energy
    mesh.add_horizontal_advection(d.spe_lnps_n, d.grid_u_c, d.grid_v_c, d.grid_δlnps)
    d.grid_δlnps -= d.grid_ps_c * d.grid_div
    mesh.trans_grid_to_spherical(d.grid_δlnps, d.spe_δlnps)
    sw_implicit_correction(integ, d.spe_div_c, d.spe_div_p, d.spe_lnps_c, d.spe_lnps_p, h_0,
                           d.spe_δdiv, d.spe_δlnps)
    integ.compute_spectral_damping(d.spe_vor_c, d.spe_vor_p, d.spe_δvor)
    integ.compute_spectral_damping(d.spe_div_c, d.spe_div_p, d.spe_δdiv)
    integ.compute_spectral_damping(d.spe_lnps_c, d.spe_lnps_p, d.spe_δlnps)
    integ.filtered_leapfrog(d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)
    integ.filtered_leapfrog(d.spe_δdiv, d.spe_div_p, d.spe_div_c, d.spe_div_n)
    integ.filtered_leapfrog(d.spe_δlnps, d.spe_lnps_p, d.spe_lnps_c, d.spe_lnps_n)
    mesh.trans_spherical_to_grid(d.spe_vor_n, d.grid_vor)
    mesh.trans_spherical_to_grid(d.spe_div_n, d.grid_div)
    mesh.trans_spherical_to_grid(d.spe_lnps_n, d.grid_ps_n)
    mesh.uv_grid_from_vor_div(d.spe_vor_n, d.spe_div_n, d.grid_u_n, d.grid_v_n)
    time_advance(d)
