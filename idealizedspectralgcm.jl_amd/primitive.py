"""Primitive-equation (Held-Suarez) mirror over the C ABI.

Reference: src/Atmos_Spectral/Vert_Coordinate.jl, Press_And_Geopot.jl,
Semi_Implicit.jl, Spectral_Dynamics.jl and src/Atmos_Param/HS_Forcing.jl.
Names follow the reference (trailing ``!`` dropped).  Column work and the
semi-implicit solve are fused kernels (jgcm_hs_*); the call order of
Spectral_Dynamics! is kept, so every intermediate the reference keeps in
Dyn_Data is produced in the same place, except the scratch
``grid_energy_full`` (the reference also uses it as a temporary for the
energy correction; here that temporary stays on the device).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call
from .mesh import (_ptr, Trans_Grid_To_Spherical, Trans_Spherical_To_Grid,
                   Trans_Grid_To_Spherical_Batch, Trans_Spherical_To_Grid_Batch, BATCH_DIVCOS,
                   BATCH_VORDIV, INV_GRAD_LAMBDA, INV_GRAD_THETA, INV_UCOS, INV_VCOS,
                   Vor_Div_From_Grid_UV, UV_Grid_From_Vor_Div)
from .dyn_data import ROTATED_STATE, Time_Advance_Tracers, _rotate_state
from .time_integrator import Get_Δt

_VC = {"even_sigma": 0, "uneven_sigma": 1, "hybrid": 2, "mcm": 3, "v197": 4}
_VD = {"simmons_and_burridge": 0}
_VA = {"second_centered_wts": 0, "second_centered": 1}


class _AtmoParams(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_double), ("omega", ctypes.c_double), ("grav", ctypes.c_double),
                ("rdgas", ctypes.c_double), ("kappa", ctypes.c_double), ("cp_air", ctypes.c_double),
                ("do_mass_correction", ctypes.c_int), ("do_energy_correction", ctypes.c_int)]


class _HSParams(ctypes.Structure):
    _fields_ = [("sigma_b", ctypes.c_double), ("k_f", ctypes.c_double), ("k_a", ctypes.c_double),
                ("k_s", ctypes.c_double), ("delta_T_y", ctypes.c_double),
                ("delta_theta_z", ctypes.c_double)]


def _atmo(a):
    if a.do_water_correction:
        raise RuntimeError("water correction has not implemented")   # Spectral_Dynamics.jl:24
    return _AtmoParams(a.radius, a.omega, a.grav, a.rdgas, a.kappa, a.cp_air,
                       int(a.do_mass_correction), int(a.do_energy_correction))


def _hs(params):
    return _HSParams(params["σ_b"], params["k_f"], params["k_a"], params["k_s"], params["ΔT_y"],
                     params["Δθ_z"])


class Vert_Coordinate:
    """``Vert_Coordinate(nλ, nθ, nd, vert_coord_option, vert_difference_option,
    vert_advect_scheme, p_ref=101325., zero_top=True, scale_heights=4.0,
    surf_res=1.0, p_press=0.1, p_sigma=0.3, exponent=2.5)``
    (Vert_Coordinate.jl:42-56).  uneven_sigma and hybrid follow the
    reference's docstring and call structure (the reference itself cannot
    run them; include/jgcm.h states the formulas).  The mesh defaults to the
    last mesh created on this (nθ, nλ) grid."""

    def __init__(self, nλ, nθ, nd, vert_coord_option, vert_difference_option,
                 vert_advect_scheme, p_ref=101325.0, zero_top=True, scale_heights=4.0,
                 surf_res=1.0, p_press=0.1, p_sigma=0.3, exponent=2.5, mesh=None):
        from .mesh import _MESHES
        if mesh is None:
            for ref in reversed(_MESHES):
                m = ref()
                if m is not None and m._h is not None and (m.nθ, m.nλ) == (nθ, nλ):
                    mesh = m
                    break
        assert mesh is not None, "create the Spectral_Spherical_Mesh first"
        for opt, table in ((vert_coord_option, _VC), (vert_difference_option, _VD),
                           (vert_advect_scheme, _VA)):
            if opt not in table:
                raise RuntimeError(f"{opt} is not a valid value for option")
        self.mesh = mesh
        self.nλ, self.nθ, self.nd = nλ, nθ, nd
        self.vert_coord_option = vert_coord_option
        self.vert_difference_option = vert_difference_option
        self.vert_advect_scheme = vert_advect_scheme
        self.p_ref, self.zero_top = float(p_ref), bool(zero_top)
        h = ctypes.c_void_p()
        call("jgcm_vert_coord_create_ex", mesh.handle, nd, _VC[vert_coord_option],
             _VD[vert_difference_option], _VA[vert_advect_scheme], self.p_ref,
             1 if zero_top else 0, float(scale_heights), float(surf_res), float(p_press),
             float(p_sigma), float(exponent), ctypes.byref(h))
        self._h = h
        ak, bk = np.zeros(nd + 1), np.zeros(nd + 1)
        call("jgcm_vert_coord_get", h, ak.ctypes.data_as(ctypes.c_void_p),
             bk.ctypes.data_as(ctypes.c_void_p))
        self.ak, self.bk = ak, bk
        self.Δak, self.Δbk = ak[1:] - ak[:-1], bk[1:] - bk[:-1]

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                call("jgcm_vert_coord_destroy", self._h)
                self._h = None
        except Exception:
            pass


def Pressure_Variables(vert_coord, grid_ps, grid_p_half, grid_Δp, grid_lnp_half, grid_p_full,
                       grid_lnp_full):
    """Press_And_Geopot.jl:36-67."""
    assert grid_ps.shape[0] == 1                                   # :41
    call("jgcm_pressure_variables", vert_coord.handle, _ptr(grid_ps), _ptr(grid_p_half),
         _ptr(grid_Δp), _ptr(grid_lnp_half), _ptr(grid_p_full), _ptr(grid_lnp_full))


def Compute_Pressures_And_Heights(atmo_data, vert_coord, grid_ps, grid_geopots, grid_t,
                                  grid_p_half, grid_Δp, grid_lnp_half, grid_p_full, grid_lnp_full,
                                  grid_z_full, grid_z_half):
    """Press_And_Geopot.jl:3-19."""
    a = _atmo(atmo_data)
    call("jgcm_compute_pressures_and_heights", vert_coord.handle, ctypes.byref(a), _ptr(grid_ps),
         _ptr(grid_geopots), _ptr(grid_t), _ptr(grid_p_half), _ptr(grid_Δp), _ptr(grid_lnp_half),
         _ptr(grid_p_full), _ptr(grid_lnp_full), _ptr(grid_z_full), _ptr(grid_z_half))


def Compute_Geopotential(vert_coord, atmo_data, grid_lnp_half, grid_lnp_full, grid_t,
                         grid_geopots, grid_geopot_full, grid_geopot_half):
    """Press_And_Geopot.jl:71-111."""
    a = _atmo(atmo_data)
    call("jgcm_compute_geopotential", vert_coord.handle, ctypes.byref(a), _ptr(grid_lnp_half),
         _ptr(grid_lnp_full), _ptr(grid_t), _ptr(grid_geopots), _ptr(grid_geopot_full),
         _ptr(grid_geopot_half))


def Vert_Advection(vert_coord, r, dz, w, Δt, vert_advect_scheme, rdt):
    """Vert_Coordinate.jl:171-210 (Δt is unused, as in the reference)."""
    assert vert_advect_scheme == vert_coord.vert_advect_scheme
    call("jgcm_vert_advection", vert_coord.handle, _ptr(r), _ptr(dz), _ptr(w), _ptr(rdt))


def Four_In_One(vert_coord, atmo_data, grid_div, grid_u, grid_v, grid_ps, grid_Δp, grid_lnp_half,
                grid_lnp_full, grid_p_full, grid_dλ_ps, grid_dθ_ps, grid_t, grid_M_half,
                grid_w_full, grid_δu, grid_δv, grid_δps, grid_δt):
    """Spectral_Dynamics.jl:88-183."""
    a = _atmo(atmo_data)
    call("jgcm_four_in_one", vert_coord.handle, ctypes.byref(a), *map(_ptr, (
        grid_div, grid_u, grid_v, grid_ps, grid_Δp, grid_lnp_half, grid_lnp_full, grid_p_full,
        grid_dλ_ps, grid_dθ_ps, grid_t, grid_M_half, grid_w_full, grid_δu, grid_δv, grid_δps,
        grid_δt)))


class Semi_Implicit_Solver:
    """``Semi_Implicit_Solver(vert_coord, atmo_data, integrator, ps_ref, t_ref,
    wave_numbers)`` (Semi_Implicit.jl:40-97)."""

    def __init__(self, vert_coord, atmo_data, integrator, ps_ref, t_ref, wave_numbers):
        self.vert_coord, self.atmo_data, self.integrator = vert_coord, atmo_data, integrator
        self.nd = vert_coord.nd
        self.ps_ref = float(ps_ref)
        self.t_ref = np.ascontiguousarray(t_ref, dtype=np.float64)
        assert self.t_ref.size == self.nd
        self.wave_numbers = wave_numbers
        self.num_wavenumbers = wave_numbers.shape[1] - 2
        a = _atmo(atmo_data)
        h = ctypes.c_void_p()
        call("jgcm_semi_implicit_create", vert_coord.handle, ctypes.byref(a), integrator.handle,
             self.ps_ref, self.t_ref.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
        self._h = h
        self._corrections_ready = False
        vert_coord.mesh._semi_implicit = self
        nd = self.nd
        self.h = self._get(0, (nd,))
        self.div_mat = self._get(1, (nd, nd))

    def _get(self, which, shape):
        out = np.zeros(int(np.prod(shape)))
        call("jgcm_semi_implicit_get", self._h, which, out.ctypes.data_as(ctypes.c_void_p), out.size)
        return out.reshape(shape)

    @property
    def wave_matrix(self):
        nd = self.nd
        return self._get(2, (self.num_wavenumbers + 1, nd, nd))

    @property
    def handle(self):
        return self._h

    def semi_implicit_update(self):
        """Update_Init_Step!(semi_implicit) (Semi_Implicit.jl:182-188)."""
        call("jgcm_semi_implicit_update_init_step", self._h)

    def scalars(self):
        out = np.zeros(7)
        call("jgcm_hs_scalars", self._h, out.ctypes.data_as(ctypes.c_void_p), 7)
        return out

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                call("jgcm_semi_implicit_destroy", self._h)
                self._h = None
        except Exception:
            pass


def Implicit_Correction_HS(semi_implicit, vert_coord, atmo_data, spe_div_c, spe_div_p,
                           spe_lnps_c, spe_lnps_p, spe_t_c, spe_t_p, spe_δdiv, spe_δlnps, spe_δt):
    """Semi_Implicit.jl:462-513 (reached through Implicit_Correction)."""
    call("jgcm_implicit_correction", semi_implicit.handle, *map(_ptr, (
        spe_div_c, spe_div_p, spe_lnps_c, spe_lnps_p, spe_t_c, spe_t_p, spe_δdiv, spe_δlnps,
        spe_δt)))


def _physics(si, d, params, dt, forcing):
    # the pressure variables of ps_c are evaluated inside the column passes and
    # not stored here: Dyn_Data's pressure fields keep the values the last
    # Pressure_Variables! / Compute_Pressures_And_Heights! wrote (HS_Forcing!
    # reads them, Spectral_Dynamics.jl:457-463, without writing them)
    P = lambda t: None if t is None else _ptr(t)
    call("jgcm_hs_physics", si.handle, ctypes.byref(_hs(params)), int(dt), *map(P, (
        d.grid_ps_c, d.grid_u_p, d.grid_v_p, d.grid_t_p, d.grid_ps_p, None, None, None, None,
        None, d.grid_δu, d.grid_δv, d.grid_t_eq, d.grid_δt, d.grid_δps)), 1 if forcing else 0)
    si._corrections_ready = True


_ZERO_HS = {"σ_b": 0.7, "k_f": 0.0, "k_a": 0.0, "k_s": 0.0, "ΔT_y": 0.0, "Δθ_z": 0.0}


def Spectral_Dynamics_Physics(atmo_data, mesh, dyn_data, Δt, physics_params, semi_implicit=None):
    """Spectral_Dynamics.jl:452-467: δps = 0 and HS_Forcing! on the p-fields.
    The fused kernel also evaluates Compute_Corrections_Init for the dynamics
    that follow (it only depends on those p-fields and tendencies)."""
    si = semi_implicit or mesh._semi_implicit
    _physics(si, dyn_data, physics_params, Δt, True)


class _StepScratch:
    """Device temporaries of one HS step that the reference keeps inside its
    mesh helpers: the grid gradients of T used by Add_Horizontal_Advection!,
    and the gradients of ln ps a step carries to the next one."""

    def __init__(self, mesh, nd):
        g = mesh.grid_zeros(2 * nd + 2)
        self.grid_dl_t, self.grid_dt_t = g[0:nd], g[nd:2 * nd]
        self.carry_dl_lnps, self.carry_dt_lnps = g[2 * nd:2 * nd + 1], g[2 * nd + 1:]


def _scratch(mesh, d):
    sc = getattr(d, "_step_scratch", None)
    if sc is None:
        sc = d._step_scratch = _StepScratch(mesh, d.spe_vor_c.shape[0])
    return sc


def Spectral_Dynamics(mesh, vert_coord, atmo_data, dyn_data, semi_implicit,
                      pressure_variables=True, physics_params=None, carry_gradients=False,
                      rotate=False):
    """Spectral_Dynamics.jl:206-365.  The transforms of the step are batched:
    the gradients of ln ps and T (:259, :293) go back to the grid in one
    launch pair with the spectral gradient operators applied on the way in,
    the tendencies δlnps, δT, Φ+KE and (δu, δv) -> (δvor, δdiv) (:280, :294,
    :304, :308) go forward in one, and the new state (:331-335, u and v through
    the u cos θ / v cos θ operators) comes back in one; the grid part of
    Add_Horizontal_Advection! for T is fused into the column pass, ∇² of the
    energy into the semi-implicit solve, and the seven damping / leapfrog
    calls (:319-328) into one launch.  Results equal the unbatched sequence.
    ``pressure_variables=False`` skips the closing Pressure_Variables! (:355)
    when the caller recomputes the same fields next (Atmosphere_Update!).
    ``physics_params`` (Atmosphere_Update!) runs Spectral_Dynamics_Physics!
    (:452-467) inside the step's column pass, its tendencies kept in
    registers (jgcm_hs_physics_column_dynamics).

    ``carry_gradients`` (with ``physics_params``): the closing inverse batch
    also forms the next step's gradients of ln ps and T (:259, :293) from the
    new state, and the next step that also carries uses them instead of
    transforming its current state again (one launch pair fewer per step).
    The current state differs from the new one only in the (0, 0)
    coefficients that Compute_Corrections! adjusts, which the gradient
    operators multiply by exact zeros, so the results are the same bit for
    bit.  Every step of a run must then carry, and nothing may change the
    state between the steps.

    ``rotate`` (with ``physics_params``): Time_Advance! (:346) renames the p,
    c, n arrays of the state (dyn_data.ROTATED_STATE) instead of copying
    them: the Robert-filtered c is written in place (as Filtered_Leapfrog!
    does), Compute_Corrections! adjusts n in place, and then p <- c, c <- n
    by name.  p and c end as the copying step leaves them, bit for bit; the
    n arrays hold stale values afterwards (the reference's n keeps the new
    state), which nothing in a step reads before writing it.  Graphs of
    rotated steps need a multiple of 3 steps (Step_Graph checks)."""
    d, si = dyn_data, semi_implicit
    integrator = si.integrator
    Δt = Get_Δt(integrator)
    sc = _scratch(mesh, d)
    h = mesh.handle
    fused = physics_params is not None
    assert fused or not carry_gradients, "carried gradients need the fused physics pass"
    assert fused or not rotate, "the rotated Time_Advance! needs the fused physics pass"
    carried = carry_gradients and getattr(d, "_carried_gradients", False)
    d._carried_gradients = False
    if not fused and not si._corrections_ready:
        # Compute_Corrections_Init + Pressure_Variables! without forcing (:250-256)
        _physics(si, d, _ZERO_HS, Δt, False)
    si._corrections_ready = False
    # the physics pass may still run on the mesh's side stream (Atmosphere_Update
    # enqueues it there): the gradient batch below does not depend on it
    call("jgcm_mesh_branch", h, 0)
    # :259 Compute_Gradients!(lnps) and the gradient half of :293 Add_Horizontal_Advection!(T)
    # (carried: formed by the previous step's closing batch)
    if not carried:
        Trans_Spherical_To_Grid_Batch(
            mesh, [d.spe_lnps_c, d.spe_lnps_c, d.spe_t_c, d.spe_t_c],
            [d.grid_dλ_ps, d.grid_dθ_ps, sc.grid_dl_t, sc.grid_dt_t],
            [BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA,
             BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA])
    call("jgcm_mesh_join", h)                  # the column pass needs both
    # :256 Pressure_Variables!(ps_c) is evaluated inside the column pass (NULL fields)
    P = lambda t: None if t is None else _ptr(t)
    cols = (d.grid_geopots, d.grid_M_half, d.grid_w_full, d.grid_δu, d.grid_δv, d.grid_δps,
            d.grid_δt, d.grid_geopot_half, d.grid_geopot_full, d.grid_δlnps, d.grid_absvor,
            d.grid_energy_full, sc.grid_dl_t, sc.grid_dt_t)
    if fused:                                                                  # :452-467, :260-307
        args = (d.grid_u_p, d.grid_v_p, d.grid_t_p, d.grid_ps_p, d.grid_t_eq, d.grid_div,
                d.grid_u_c, d.grid_v_c, d.grid_ps_c, d.grid_dλ_ps, d.grid_dθ_ps, d.grid_t_c,
                d.grid_vor) + cols
        # with the p = c half of :346 Time_Advance! for the grid u, v, T, ps
        carry_in = (sc.carry_dl_lnps, sc.carry_dt_lnps) if carried else (None, None)
        call("jgcm_hs_physics_column_dynamics_ex", si.handle, ctypes.byref(_hs(physics_params)),
             int(Δt), *map(P, args + carry_in), 0 if rotate else 1)
    else:
        call("jgcm_hs_column_dynamics", si.handle, *map(P, (                   # :260-307
            d.grid_div, d.grid_u_c, d.grid_v_c, d.grid_ps_c, None, None,
            None, None, d.grid_dλ_ps, d.grid_dθ_ps, d.grid_t_c, d.grid_vor) + cols))
    # :280, :294, :304, :308
    Trans_Grid_To_Spherical_Batch(
        mesh, [d.grid_δlnps, d.grid_δt, d.grid_energy_full, d.grid_δu, d.grid_δv],
        [d.spe_δlnps, d.spe_δt, d.spe_energy, d.spe_δvor, d.spe_δdiv],
        [0, 0, 0, BATCH_VORDIV, 0])
    # :309-316 ∇²(energy), δdiv -= ∇²(energy), Implicit_Correction!, and in the same
    # pass :319-328 Compute_Spectral_Damping! (vor, div, T) and Filtered_Leapfrog!
    # (vor, div, lnps, T)
    # and the spectral part of :346 Time_Advance! (p = the filtered c, c = n)
    if rotate:   # the filtered c in place (Filtered_Leapfrog!), renamed to p below
        call("jgcm_hs_energy_implicit_leapfrog", si.handle, *map(_ptr, (
            d.spe_energy, d.spe_div_c, d.spe_div_p, d.spe_lnps_c, d.spe_lnps_p, d.spe_t_c,
            d.spe_t_p, d.spe_δdiv, d.spe_δlnps, d.spe_δt, d.spe_δvor, d.spe_vor_p, d.spe_vor_c,
            d.spe_vor_n, d.spe_div_n, d.spe_lnps_n, d.spe_t_n)))
    else:
        P3 = ctypes.c_void_p * 3
        lev = lambda f: P3(*(getattr(d, f + s).data_ptr() for s in ("_p", "_c", "_n")))
        call("jgcm_hs_implicit_leapfrog_advance", si.handle, *map(_ptr, (
            d.spe_energy, d.spe_δdiv, d.spe_δlnps, d.spe_δt, d.spe_δvor)),
            lev("spe_vor"), lev("spe_div"), lev("spe_lnps"), lev("spe_t"))
    # :331-335: vor, div, (u, v) = UV_Grid_From_Vor_Div!, T, ln ps back to the grid
    # (carrying: with the next step's gradients of ln ps and T)
    snms = [d.spe_vor_n, d.spe_div_n, d.spe_vor_n, d.spe_vor_n, d.spe_t_n, d.spe_lnps_n]
    grids = [d.grid_vor, d.grid_div, d.grid_u_n, d.grid_v_n, d.grid_t_n, d.grid_lnps]
    flags = [0, 0, BATCH_DIVCOS | INV_UCOS, BATCH_DIVCOS | INV_VCOS, 0, 0]
    if carry_gradients:
        snms += [d.spe_lnps_n, d.spe_lnps_n, d.spe_t_n, d.spe_t_n]
        grids += [sc.carry_dl_lnps, sc.carry_dt_lnps, sc.grid_dl_t, sc.grid_dt_t]
        flags += [BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA,
                  BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA]
    Trans_Spherical_To_Grid_Batch(mesh, snms, grids, flags,
                                  snms_b=[None, None, d.spe_div_n, d.spe_div_n] +
                                  [None] * (len(snms) - 4))
    if rotate:
        # :336-343 Compute_Corrections! on n in place, :346 Time_Advance! by renaming
        call("jgcm_hs_corrections", si.handle, *map(_ptr, (
            d.grid_lnps, d.grid_ps_n, d.grid_u_n, d.grid_v_n, d.grid_t_n, d.spe_lnps_n,
            d.spe_t_n)))
        _rotate_state(d)
    else:
        # :336-343 Compute_Corrections! and :346 Time_Advance!, the corrections applied
        # to the new fields while they are copied (one launch after the sums)
        names = ROTATED_STATE
        arrs = [[getattr(d, f + s) for f in names] for s in ("_p", "_c", "_n")]
        ptrs = [(ctypes.c_void_p * 8)(*[t.data_ptr() for t in ts]) for ts in arrs]
        # p = c done above for the spectral fields and (fused) the grid fields: c = n and
        # the corrections remain
        for k in range(8 if fused else 4):
            ptrs[0][k] = None
        counts = (ctypes.c_size_t * 8)(*[t.numel() * t.element_size() // 8 for t in arrs[1]])
        call("jgcm_hs_corrections_advance", si.handle, _ptr(d.grid_lnps), *ptrs, counts)
    Time_Advance_Tracers(d)
    d._carried_gradients = carry_gradients
    if pressure_variables:
        Pressure_Variables(vert_coord, d.grid_ps_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half,
                           d.grid_p_full, d.grid_lnp_full)                        # :355


def Atmosphere_Update(mesh, atmo_data, vert_coord, semi_implicit, dyn_data, physics_params,
                      overlap=False, fused=True, eager_heights=False, carry_gradients=False,
                      rotate=False):
    """Spectral_Dynamics.jl:470-490.  By default (``fused``) the physics pass
    (HS forcing on the p-fields, the correction partials) runs inside the
    step's column pass, after the gradient transforms, its tendencies kept in
    registers.  ``fused=False`` runs it as its own pass first; with
    ``overlap`` as well, that pass goes to the mesh's side stream beside the
    first transforms (parallel branches of a captured step).  The results
    agree to rounding.  On MI355X at T85 L20 (one-step graphs): fused 204 us,
    separate serial 210 us, separate overlapped 218 us (the branch's fork and
    join cost more than the pass).  The closing Compute_Pressures_And_Heights!
    (:485) is deferred to the first read of its outputs (Dyn_Data.HEIGHT_FIELDS;
    nothing in the next step reads them); ``eager_heights`` runs it here, in
    the reference's order.  ``carry_gradients`` (fused only): each step's
    closing inverse batch also forms the next step's gradients of ln ps and T
    (Spectral_Dynamics); for runs whose steps all carry and that do not touch
    the state between steps (driver loops, captured multi-step graphs).
    ``rotate`` (fused only): Time_Advance! by renaming the state's p, c, n
    arrays instead of copying them (Spectral_Dynamics); the n arrays are
    stale after the step, and graphs of rotated steps take a multiple of 3
    steps."""
    Δt = Get_Δt(semi_implicit.integrator)
    assert not carry_gradients or (fused and not overlap), "carried gradients need the fused step"
    assert not rotate or (fused and not overlap), "the rotated Time_Advance! needs the fused step"
    if fused and not overlap:
        Spectral_Dynamics(mesh, vert_coord, atmo_data, dyn_data, semi_implicit,
                          pressure_variables=False, physics_params=physics_params,
                          carry_gradients=carry_gradients, rotate=rotate)
    else:
        if overlap:
            call("jgcm_mesh_branch", mesh.handle, 1)
        Spectral_Dynamics_Physics(atmo_data, mesh, dyn_data, Δt, physics_params, semi_implicit)
        # the closing Pressure_Variables! of Spectral_Dynamics! (:355) is recomputed
        # identically by Compute_Pressures_And_Heights! below
        Spectral_Dynamics(mesh, vert_coord, atmo_data, dyn_data, semi_implicit,
                          pressure_variables=False)
    d = dyn_data
    if eager_heights:
        d._heights_src = None
        Compute_Pressures_And_Heights(atmo_data, vert_coord, d.grid_ps_c, d.grid_geopots,
                                      d.grid_t_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half,
                                      d.grid_p_full, d.grid_lnp_full, d.grid_z_full,
                                      d.grid_z_half)
    else:  # :485 computed when first read (Dyn_Data.HEIGHT_FIELDS)
        d._defer_heights(mesh, atmo_data, vert_coord)


def Get_Topography(grid_geopots):
    """Spectral_Dynamics.jl:367-370."""
    grid_geopots.zero_()


def Spectral_Initialize_Fields(mesh, atmo_data, vert_coord, sea_level_ps_ref, init_t,
                               grid_geopots, dyn_data, grid_t_noise=None):
    """Spectral_Dynamics.jl:372-449.  ``grid_t_noise`` (optional device grid
    field) is added to the isothermal temperature before the spectral
    filtering (the seeded-noise Held-Suarez benchmark state)."""
    dyn_data._carried_gradients = False   # a new state: nothing carried
    d = dyn_data
    nd = mesh.nd
    G = d.grid_t_c.numel()
    call("jgcm_fill", mesh.handle, _ptr(d.grid_t_c), float(init_t), G)
    if grid_t_noise is not None:
        call("jgcm_axpby", mesh.handle, G, 1.0, _ptr(d.grid_t_c), 1.0, _ptr(grid_t_noise),
             _ptr(d.grid_t_c))
    g1 = d.grid_lnps.numel()
    call("jgcm_map", mesh.handle, 2, g1, _ptr(grid_geopots), -1.0 / (atmo_data.rdgas * init_t),
         float(np.log(sea_level_ps_ref)), _ptr(d.grid_lnps))
    call("jgcm_map", mesh.handle, 0, g1, _ptr(d.grid_lnps), 0.0, 0.0, _ptr(d.grid_ps_c))
    d.spe_div_c.zero_()
    d.spe_vor_c.zero_()
    pert = 1.0e-7 / np.sqrt(2.0)
    mesh.synchronize()
    for k in range(nd - 3, nd):
        for (n, m) in ((4, 1), (8, 5), (3, 1), (7, 5)):
            d.spe_vor_c[k, n, m] = pert
    torch.cuda.synchronize()
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    Trans_Grid_To_Spherical(mesh, d.grid_t_c, d.spe_t_c)
    Trans_Spherical_To_Grid(mesh, d.spe_t_c, d.grid_t_c)
    Trans_Grid_To_Spherical(mesh, d.grid_lnps, d.spe_lnps_c)
    Trans_Spherical_To_Grid(mesh, d.spe_lnps_c, d.grid_lnps)
    call("jgcm_map", mesh.handle, 0, g1, _ptr(d.grid_lnps), 0.0, 0.0, _ptr(d.grid_ps_c))
    Vor_Div_From_Grid_UV(mesh, d.grid_u_c, d.grid_v_c, d.spe_vor_c, d.spe_div_c)
    UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    Trans_Spherical_To_Grid(mesh, d.spe_div_c, d.grid_div)
    Pressure_Variables(vert_coord, d.grid_ps_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half,
                       d.grid_p_full, d.grid_lnp_full)
    for f in ("vor", "div", "lnps", "t"):
        src, dst = getattr(d, f"spe_{f}_c"), getattr(d, f"spe_{f}_p")
        call("jgcm_copy", mesh.handle, _ptr(dst), _ptr(src), src.numel() * src.element_size())
    for f in ("u", "v", "ps", "t"):
        src, dst = getattr(d, f"grid_{f}_c"), getattr(d, f"grid_{f}_p")
        call("jgcm_copy", mesh.handle, _ptr(dst), _ptr(src), src.numel() * src.element_size())
