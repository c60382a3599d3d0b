"""Barotropic and shallow-water steppers, composed from the C ABI exactly as
the reference composes them from Julia functions.

Reference: src/Atmos_Spectral/Barotropic_Dynamics.jl,
src/Atmos_Spectral/Shallow_Water_Dynamics.jl.  Every array operation is a
library call on the mesh stream (so a whole step can be captured into a HIP
graph); no arithmetic happens on the host.  The per-step ``@show norm(...)``
diagnostics of the reference are omitted (``verbose=True`` prints them, which
synchronises the stream).
"""
import ctypes

import numpy as np

from ._lib import call
from .dyn_data import (BAROTROPIC_STATE, SW_STATE, Time_Advance_Model, Time_Advance_Tracers,
                       _rotate_state)
from .mesh import (_ptr, Vor_Div_From_Grid_UV, Apply_Laplacian, Add_Horizontal_Advection,
                   Trans_Spherical_To_Grid_Batch, Trans_Grid_To_Spherical_Batch, BATCH_DIVCOS,
                   BATCH_VORDIV, INV_UCOS, INV_VCOS, INV_GRAD_LAMBDA, INV_GRAD_THETA)
from .time_integrator import Damped_Leapfrog, Damped_Leapfrog_Batch


def _show(*ts):
    import torch
    print(tuple(float(torch.linalg.norm(t)) for t in ts))


def Barotropic_Dynamics(mesh, atmo_data, dyn_data, integrator, verbose=False):
    """Barotropic_Dynamics.jl:2-45.  Compute_Spectral_Damping! and
    Filtered_Leapfrog! run as one kernel, and the three inverse transforms of
    the new vorticity (ζ, u, v: :38-40) as one batched launch pair; the
    results equal the separate calls bit for bit."""
    d = dyn_data
    n = mesh._nlev(d.grid_vor, "grid")
    # grid_absvor = vor + f;  δu = absvor v;  δv = -absvor u        (:26-29)
    call("jgcm_vorticity_flux", mesh.handle, _ptr(d.grid_vor), _ptr(d.grid_u_c), _ptr(d.grid_v_c),
         atmo_data.omega, _ptr(d.grid_absvor), _ptr(d.grid_δu), _ptr(d.grid_δv), n, 0)
    Vor_Div_From_Grid_UV(mesh, d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    Damped_Leapfrog(integrator, d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)  # :34-36
    # :38-40 Trans_Spherical_To_Grid!(vor) and UV_Grid_From_Vor_Div!(vor, 0)
    Trans_Spherical_To_Grid_Batch(
        mesh, [d.spe_vor_n, d.spe_vor_n, d.spe_vor_n], [d.grid_vor, d.grid_u_n, d.grid_v_n],
        [0, BATCH_DIVCOS | INV_UCOS, BATCH_DIVCOS | INV_VCOS],
        snms_b=[None, d.spe_zeros, d.spe_zeros])
    if verbose:
        _show(d.grid_u_n, d.grid_v_n, d.grid_vor)
    Time_Advance_Model(d, BAROTROPIC_STATE)  # :44 (the arrays this step writes)


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
    """Shallow_Water_Dynamics.jl:88-179.  The four forward transforms (δu, δv
    -> δζ, δD at :133; KE at :140; δh at :149) run as one batched launch pair
    once their grid fields exist (the height tendency is formed first, it
    does not depend on δu, δv); the three dampings and leapfrogs (:159-165)
    as one kernel; and the inverse transforms of the new state (ζ, D, h, and
    u, v from ζ, D: :169-172) as one batched launch pair.  Each operation
    keeps its arithmetic, so the results equal the call-by-call sequence bit
    for bit."""
    d = dyn_data
    n = mesh._nlev(d.grid_vor, "grid")
    # abs vor and explicit residual δu += absvor v, δv += -absvor u    (:126-130)
    call("jgcm_vorticity_flux", mesh.handle, _ptr(d.grid_vor), _ptr(d.grid_u_c), _ptr(d.grid_v_c),
         atmo_data.omega, _ptr(d.grid_absvor), _ptr(d.grid_δu), _ptr(d.grid_δv), n, 1)
    # kinetic energy on the grid                                        (:139)
    call("jgcm_kinetic_energy", mesh.handle, _ptr(d.grid_u_c), _ptr(d.grid_v_c), None,
         _ptr(d.grid_d_full1), n)
    _sw_dynamics_rest(mesh, h_0, d, integrator, verbose)


def Shallow_Water_Step(mesh, atmo_data, h_0, dyn_data, integrator, kappa_m, kappa_t, h_eq,
                       carry_gradients=False, rotate=False):
    """Shallow_Water_Physics! then Shallow_Water_Dynamics! (the drivers' step,
    Shallow_Water_Dynamics.jl:7-21 and :88-179) in fused passes (no reference
    counterpart): the physics, the absolute-vorticity residual, the kinetic
    energy and p = c of the grid state in one grid pass (jgcm_sw_explicit);
    the advection and h div terms of the height tendency in one
    (jgcm_sw_height_tendency); the spectral energy, implicit correction,
    damped leapfrog and spectral Time_Advance! in one (jgcm_sw_spectral_step);
    c = n of the grid state as one batched copy.  Equal to the two calls bit
    for bit.

    ``carry_gradients``: the closing inverse batch also forms the gradient of
    the new h, which is the h whose gradient the next step's height tendency
    takes (:147), and the next step that also carries uses it instead of two
    more inverse transforms.  Same results bit for bit; every step of the run
    must carry, and nothing may change the state between the steps.

    ``rotate``: Time_Advance! by renaming the p, c, n arrays of the state
    (dyn_data.SW_STATE) instead of copying them: the grid pass skips p = c,
    the spectral pass writes the filtered c in place, the closing copy goes.
    p and c end bit for bit as the copying step leaves them; the n arrays
    hold stale values after the step, and graphs of rotated steps take a
    multiple of 3 steps (Step_Graph checks)."""
    d = dyn_data
    carried = carry_gradients and getattr(d, "_carried_sw", False)
    d._carried_sw = False
    if carry_gradients and getattr(d, "_sw_carry", None) is None:
        d._sw_carry = (mesh.grid_zeros(d.grid_u_c.shape[0]),
                       mesh.grid_zeros(d.grid_u_c.shape[0]))   # before any capture
    sw_grid_pass(mesh, atmo_data, d, kappa_m, kappa_t, h_eq, carried, advance_p=not rotate)
    Trans_Grid_To_Spherical_Batch(mesh, *sw_forward_batch(d))
    sw_spectral_pass(integrator, h_0, d, advance=not rotate)
    snms, grids, flags, snms_b = sw_inverse_batch(d, carry_gradients)
    Trans_Spherical_To_Grid_Batch(mesh, snms, grids, flags, snms_b=snms_b)
    if rotate:
        Time_Advance_Tracers(d)
        _rotate_state(d, SW_STATE)
    else:
        sw_closing_copy(mesh, d)
    d._carried_sw = carry_gradients


# The stages of Shallow_Water_Step, also run stage by stage over several
# meshes by the mixed-truncation ensemble (one transform plan per direction).
def sw_grid_pass(mesh, atmo_data, d, kappa_m, kappa_t, h_eq, carried, advance_p=True):
    """The explicit grid pass and the height tendency's grid terms (:147-148);
    `carried`: the gradient of h comes from the previous step's inverse batch;
    `advance_p`: the pass also sets p = c for the grid u, v, h."""
    n = mesh._nlev(d.grid_vor, "grid")
    P = (lambda t: _ptr(t)) if advance_p else (lambda t: None)
    call("jgcm_sw_explicit", mesh.handle, _ptr(d.grid_u_c), _ptr(d.grid_v_c), _ptr(d.grid_ps_c),
         _ptr(h_eq), float(kappa_m), float(kappa_t), _ptr(d.grid_vor), atmo_data.omega,
         _ptr(d.grid_δu), _ptr(d.grid_δv), _ptr(d.grid_δlnps), _ptr(d.grid_absvor),
         _ptr(d.grid_d_full1), P(d.grid_u_p), P(d.grid_v_p), P(d.grid_ps_p), n)
    if carried:
        call("jgcm_sw_height_tendency_carried", mesh.handle, _ptr(d._sw_carry[0]),
             _ptr(d._sw_carry[1]), _ptr(d.grid_u_c), _ptr(d.grid_v_c), _ptr(d.grid_ps_c),
             _ptr(d.grid_div), _ptr(d.grid_δlnps), n)
    else:
        call("jgcm_sw_height_tendency", mesh.handle, _ptr(d.spe_lnps_n), _ptr(d.grid_u_c),
             _ptr(d.grid_v_c), _ptr(d.grid_ps_c), _ptr(d.grid_div), _ptr(d.grid_δlnps), n)


def sw_forward_batch(d):
    """(grid arrays, spectral arrays, flags) of the step's forward batch:
    :133 Vor_Div_From_Grid_UV!, :140 and :149 Trans_Grid_To_Spherical!."""
    return ([d.grid_δu, d.grid_δv, d.grid_d_full1, d.grid_δlnps],
            [d.spe_δvor, d.spe_δdiv, d.spe_d1, d.spe_δlnps], [BATCH_VORDIV, 0, 0, 0])


def sw_spectral_pass(integrator, h_0, d, advance=True):
    """:141-143, :150-165 and (`advance`) the spectral part of :177 in one
    kernel; without `advance` the filtered c is written in place."""
    n = d.spe_δvor.shape[0]
    P3 = ctypes.c_void_p * 3
    lev = lambda f: P3(*(getattr(d, f + s).data_ptr() for s in ("_p", "_c", "_n")))
    call("jgcm_sw_spectral_step_ex", integrator.handle, float(h_0), _ptr(d.spe_d1),
         _ptr(d.spe_energy), _ptr(d.spe_δvor), _ptr(d.spe_δdiv), _ptr(d.spe_δlnps),
         lev("spe_vor"), lev("spe_div"), lev("spe_lnps"), n, 1 if advance else 0)


def sw_inverse_batch(d, carry_gradients):
    """(spectral arrays, grid arrays, flags, second operands) of the step's
    closing inverse batch: :169-172 the new state on the grid (carrying: and
    the gradient of the new h)."""
    snms = [d.spe_vor_n, d.spe_div_n, d.spe_lnps_n, d.spe_vor_n, d.spe_vor_n]
    grids = [d.grid_vor, d.grid_div, d.grid_ps_n, d.grid_u_n, d.grid_v_n]
    flags = [0, 0, 0, BATCH_DIVCOS | INV_UCOS, BATCH_DIVCOS | INV_VCOS]
    snms_b = [None, None, None, d.spe_div_n, d.spe_div_n]
    if carry_gradients:
        snms += [d.spe_lnps_n, d.spe_lnps_n]
        grids += [d._sw_carry[0], d._sw_carry[1]]
        flags += [BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA]
        snms_b += [None, None]
    return snms, grids, flags, snms_b


def sw_closing_copy(mesh, d):
    """:177 the c = n half of Time_Advance! for the grid state, then the tracers."""
    fields = [(d.grid_u_c, d.grid_u_n), (d.grid_v_c, d.grid_v_n), (d.grid_ps_c, d.grid_ps_n)]
    Pn = ctypes.c_void_p * len(fields)
    nbytes = (ctypes.c_size_t * len(fields))(*(c.numel() * c.element_size() for c, _ in fields))
    call("jgcm_copy_batch", mesh.handle, len(fields), Pn(*(c.data_ptr() for c, _ in fields)),
         Pn(*(x.data_ptr() for _, x in fields)), nbytes)
    Time_Advance_Tracers(d)


def _sw_dynamics_rest(mesh, h_0, d, integrator, verbose):
    """Shallow_Water_Dynamics.jl:133-177 once the explicit residual and the
    kinetic energy (in grid_d_full1) are on the grid."""
    spec = d.spe_energy.numel() * 2
    grid = d.grid_u_c.numel()
    grid_kin, spe_kin = d.grid_d_full1, d.spe_d1
    # height tendency on the grid                                       (:147-148)
    Add_Horizontal_Advection(mesh, d.spe_lnps_n, d.grid_u_c, d.grid_v_c, d.grid_δlnps)
    call("jgcm_xmy", mesh.handle, grid, -1.0, _ptr(d.grid_ps_c), _ptr(d.grid_div),
         _ptr(d.grid_δlnps), 1)
    # :133 Vor_Div_From_Grid_UV!, :140 and :149 Trans_Grid_To_Spherical!
    Trans_Grid_To_Spherical_Batch(
        mesh, [d.grid_δu, d.grid_δv, grid_kin, d.grid_δlnps],
        [d.spe_δvor, d.spe_δdiv, spe_kin, d.spe_δlnps], [BATCH_VORDIV, 0, 0, 0])
    # δdiv -= ∇²(KE + h)                                                (:141-143)
    call("jgcm_axpby", mesh.handle, spec, 1.0, _ptr(spe_kin), 1.0, _ptr(d.spe_lnps_c),
         _ptr(d.spe_energy))
    Apply_Laplacian(mesh, d.spe_energy)
    call("jgcm_axpby", mesh.handle, spec, 1.0, _ptr(d.spe_δdiv), -1.0, _ptr(d.spe_energy),
         _ptr(d.spe_δdiv))
    Implicit_Correction(integrator, d.spe_div_c, d.spe_div_p, d.spe_lnps_c, d.spe_lnps_p, h_0,
                        d.spe_δdiv, d.spe_δlnps)
    # :159-165 Compute_Spectral_Damping! and Filtered_Leapfrog! of ζ, D, h
    Damped_Leapfrog_Batch(integrator, [
        (d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n, True),
        (d.spe_δdiv, d.spe_div_p, d.spe_div_c, d.spe_div_n, True),
        (d.spe_δlnps, d.spe_lnps_p, d.spe_lnps_c, d.spe_lnps_n, True)])
    # :169-172 the new state on the grid
    Trans_Spherical_To_Grid_Batch(
        mesh, [d.spe_vor_n, d.spe_div_n, d.spe_lnps_n, d.spe_vor_n, d.spe_vor_n],
        [d.grid_vor, d.grid_div, d.grid_ps_n, d.grid_u_n, d.grid_v_n],
        [0, 0, 0, BATCH_DIVCOS | INV_UCOS, BATCH_DIVCOS | INV_VCOS],
        snms_b=[None, None, None, d.spe_div_n, d.spe_div_n])
    if verbose:
        _show(d.grid_u_n, d.grid_v_n, d.grid_ps_n)
    Time_Advance_Model(d, SW_STATE)  # :177 (the arrays this step writes)
