"""Barotropic and shallow-water steppers, composed from the C ABI exactly as
the reference composes them from Julia functions.

Reference: src/Atmos_Spectral/Barotropic_Dynamics.jl,
src/Atmos_Spectral/Shallow_Water_Dynamics.jl.  Every array operation is a
library call on the mesh stream (so a whole step can be captured into a HIP
graph); no arithmetic happens on the host.  The per-step ``@show norm(...)``
diagnostics of the reference are omitted (``verbose=True`` prints them, which
synchronises the stream).
"""
import numpy as np

from ._lib import call
from .dyn_data import Time_Advance
from .mesh import (_ptr, Vor_Div_From_Grid_UV, Trans_Spherical_To_Grid, Trans_Grid_To_Spherical,
                   UV_Grid_From_Vor_Div, Apply_Laplacian, Add_Horizontal_Advection)
from .time_integrator import Compute_Spectral_Damping, Filtered_Leapfrog_step


def _show(*ts):
    import torch
    print(tuple(float(torch.linalg.norm(t)) for t in ts))


def Barotropic_Dynamics(mesh, atmo_data, dyn_data, integrator, verbose=False):
    """Barotropic_Dynamics.jl:2-45."""
    d = dyn_data
    n = mesh._nlev(d.grid_vor, "grid")
    # grid_absvor = vor + f;  δu = absvor v;  δv = -absvor u        (:26-29)
    call("jgcm_vorticity_flux", mesh.handle, _ptr(d.grid_vor), _ptr(d.grid_u_c), _ptr(d.grid_v_c),
         atmo_data.omega, _ptr(d.grid_absvor), _ptr(d.grid_δu), _ptr(d.grid_δv), n, 0)
    Vor_Div_From_Grid_UV(mesh, d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    Compute_Spectral_Damping(integrator, d.spe_vor_c, d.spe_vor_p, d.spe_δvor)
    Filtered_Leapfrog_step(integrator, d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)
    Trans_Spherical_To_Grid(mesh, d.spe_vor_n, d.grid_vor)
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_n, d.spe_zeros, d.grid_u_n, d.grid_v_n)
    if verbose:
        _show(d.grid_u_n, d.grid_v_n, d.grid_vor)
    Time_Advance(d)


def Shallow_Water_Physics(dyn_data, kappa_m, kappa_t, h_eq):
    """Shallow_Water_Dynamics.jl:7-21 (h is stored in grid_ps_c, δh in grid_δlnps)."""
    d = dyn_data
    mesh = d.mesh
    n = mesh._nlev(d.grid_u_c, "grid")
    call("jgcm_shallow_water_physics", mesh.handle, _ptr(d.grid_u_c), _ptr(d.grid_v_c),
         _ptr(d.grid_ps_c), _ptr(h_eq), float(kappa_m), float(kappa_t), _ptr(d.grid_δu),
         _ptr(d.grid_δv), _ptr(d.grid_δlnps), n)


def Implicit_Correction(*args):
    """Two overloads, dispatched on the first argument like Julia:
    (Filtered_Leapfrog, div_c, div_p, h_c, h_p, h_0, δdiv, δh) -> the shallow-water
    correction (Shallow_Water_Dynamics.jl:25-54); (Semi_Implicit_Solver, ...) ->
    the primitive-equation correction (Semi_Implicit.jl:462-513)."""
    first = args[0]
    if hasattr(first, "semi_implicit_update"):
        from .primitive import Implicit_Correction_HS
        return Implicit_Correction_HS(*args)
    integrator, div_c, div_p, h_c, h_p, h_0, δdiv, δh = args
    mesh = integrator.mesh
    n = mesh._nlev(div_c, "spec", True)
    call("jgcm_sw_implicit_correction", integrator.handle, _ptr(div_c), _ptr(div_p), _ptr(h_c),
         _ptr(h_p), float(h_0), _ptr(δdiv), _ptr(δh), n)


def Shallow_Water_Dynamics(mesh, atmo_data, h_0, dyn_data, integrator, verbose=False):
    """Shallow_Water_Dynamics.jl:88-179."""
    d = dyn_data
    n = mesh._nlev(d.grid_vor, "grid")
    spec = d.spe_energy.numel() * 2
    grid = d.grid_u_c.numel()
    # abs vor and explicit residual δu += absvor v, δv += -absvor u    (:126-130)
    call("jgcm_vorticity_flux", mesh.handle, _ptr(d.grid_vor), _ptr(d.grid_u_c), _ptr(d.grid_v_c),
         atmo_data.omega, _ptr(d.grid_absvor), _ptr(d.grid_δu), _ptr(d.grid_δv), n, 1)
    Vor_Div_From_Grid_UV(mesh, d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    # kinetic energy, δdiv -= ∇²(KE + h)                                (:137-143)
    grid_kin, spe_kin = d.grid_d_full1, d.spe_d1
    call("jgcm_kinetic_energy", mesh.handle, _ptr(d.grid_u_c), _ptr(d.grid_v_c), None,
         _ptr(grid_kin), n)
    Trans_Grid_To_Spherical(mesh, grid_kin, spe_kin)
    call("jgcm_axpby", mesh.handle, spec, 1.0, _ptr(spe_kin), 1.0, _ptr(d.spe_lnps_c),
         _ptr(d.spe_energy))
    Apply_Laplacian(mesh, d.spe_energy)
    call("jgcm_axpby", mesh.handle, spec, 1.0, _ptr(d.spe_δdiv), -1.0, _ptr(d.spe_energy),
         _ptr(d.spe_δdiv))
    # heights                                                           (:147-149)
    Add_Horizontal_Advection(mesh, d.spe_lnps_n, d.grid_u_c, d.grid_v_c, d.grid_δlnps)
    call("jgcm_xmy", mesh.handle, grid, -1.0, _ptr(d.grid_ps_c), _ptr(d.grid_div),
         _ptr(d.grid_δlnps), 1)
    Trans_Grid_To_Spherical(mesh, d.grid_δlnps, d.spe_δlnps)
    Implicit_Correction(integrator, d.spe_div_c, d.spe_div_p, d.spe_lnps_c, d.spe_lnps_p, h_0,
                        d.spe_δdiv, d.spe_δlnps)
    Compute_Spectral_Damping(integrator, d.spe_vor_c, d.spe_vor_p, d.spe_δvor)
    Compute_Spectral_Damping(integrator, d.spe_div_c, d.spe_div_p, d.spe_δdiv)
    Compute_Spectral_Damping(integrator, d.spe_lnps_c, d.spe_lnps_p, d.spe_δlnps)
    Filtered_Leapfrog_step(integrator, d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)
    Filtered_Leapfrog_step(integrator, d.spe_δdiv, d.spe_div_p, d.spe_div_c, d.spe_div_n)
    Filtered_Leapfrog_step(integrator, d.spe_δlnps, d.spe_lnps_p, d.spe_lnps_c, d.spe_lnps_n)
    Trans_Spherical_To_Grid(mesh, d.spe_vor_n, d.grid_vor)
    Trans_Spherical_To_Grid(mesh, d.spe_div_n, d.grid_div)
    Trans_Spherical_To_Grid(mesh, d.spe_lnps_n, d.grid_ps_n)
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_n, d.spe_div_n, d.grid_u_n, d.grid_v_n)
    if verbose:
        _show(d.grid_u_n, d.grid_v_n, d.grid_ps_n)
    Time_Advance(d)
