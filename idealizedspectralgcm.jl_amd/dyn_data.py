"""Atmo_Data and Dyn_Data mirrors (device-resident state).

Reference: src/Atmos_Spectral/Atmo_Data.jl, src/Atmos_Spectral/Dyn_Data.jl.
Every field is a device tensor in the reference layout; ``Time_Advance`` is
done with ordered device copies on the mesh stream, with the reference's
semantics (p <- c, c <- n, the ``_n`` fields keep their values,
Dyn_Data.jl:326-364).
"""
import ctypes

import numpy as np
import torch

from ._lib import call
from .mesh import _ptr, current_mesh


class Atmo_Data:
    """``Atmo_Data(name, nλ, nθ, nd, do_mass_correction, do_energy_correction,
    do_water_correction, use_virtual_temperature, sinθ, radius, omega=7.292e-5,
    grav=9.80, rdgas=287.04, kappa=2/7, rvgas=461.50)`` (Atmo_Data.jl:34-55)."""

    def __init__(self, name, nλ, nθ, nd, do_mass_correction, do_energy_correction,
                 do_water_correction, use_virtual_temperature, sinθ, radius, omega=7.292e-5,
                 grav=9.80, rdgas=287.04, kappa=2.0 / 7.0, rvgas=461.50):
        self.name = name
        self.nλ, self.nθ, self.nd = nλ, nθ, nd
        self.do_mass_correction = bool(do_mass_correction)
        self.do_energy_correction = bool(do_energy_correction)
        self.do_water_correction = bool(do_water_correction)
        self.use_virtual_temperature = bool(use_virtual_temperature)
        self.radius, self.omega, self.grav = float(radius), float(omega), float(grav)
        self.rdgas, self.kappa, self.rvgas = float(rdgas), float(kappa), float(rvgas)
        self.coriolis = 2 * self.omega * np.asarray(sinθ)                   # :45
        self.cp_air = self.rdgas / self.kappa                                # :46


def Compute_Abs_Vor(mesh, grid_vor, atmo_data, grid_absvor):
    """Atmo_Data.jl:59-65 (takes the mesh and Atmo_Data in place of the
    coriolis array: f = 2Ω sinθ is formed on the device)."""
    n = mesh._nlev(grid_vor, "grid", True)
    call("jgcm_vorticity_flux", mesh.handle, _ptr(grid_vor), None, None, atmo_data.omega,
         _ptr(grid_absvor), None, None, n, 0)


_SPEC_ND = ["spe_vor_n", "spe_vor_c", "spe_vor_p", "spe_div_n", "spe_div_c", "spe_div_p",
            "spe_t_n", "spe_t_c", "spe_t_p", "spe_δvor", "spe_δdiv", "spec_δu", "spec_δv",
            "spe_δt", "spe_energy", "spe_d1", "spe_d2", "spe_zeros"]
_SPEC_1 = ["spe_lnps_n", "spe_lnps_c", "spe_lnps_p", "spe_δlnps"]
_GRID_ND = ["grid_u_n", "grid_u_c", "grid_u_p", "grid_v_n", "grid_v_c", "grid_v_p",
            "grid_t_n", "grid_t_c", "grid_t_p", "grid_vor", "grid_δvor", "grid_div", "grid_δdiv",
            "grid_δu", "grid_δv", "grid_p_full", "grid_lnp_full", "grid_Δp", "grid_δt",
            "grid_absvor", "grid_w_full", "grid_energy_full", "grid_geopot_full", "grid_z_full",
            "grid_t_eq", "grid_d_full1", "grid_d_full2"]
_GRID_1 = ["grid_ps_n", "grid_ps_c", "grid_ps_p", "grid_lnps", "grid_δlnps", "grid_δps",
           "grid_dλ_ps", "grid_dθ_ps", "grid_geopots"]
_GRID_HALF = ["grid_p_half", "grid_lnp_half", "grid_M_half", "grid_geopot_half", "grid_z_half",
              "grid_d_half1", "grid_d_half2"]


# Outputs of Compute_Pressures_And_Heights! (Press_And_Geopot.jl:3-19) at the
# end of Atmosphere_Update! (Spectral_Dynamics.jl:485), in its argument order.
# No later part of a step reads them (the column pass evaluates the pressures
# of ps itself), so Atmosphere_Update defers them: they are computed from the
# state when first read after a step (see Dyn_Data._defer_heights).
HEIGHT_FIELDS = ("grid_p_half", "grid_Δp", "grid_lnp_half", "grid_p_full", "grid_lnp_full",
                 "grid_z_full", "grid_z_half")


def _deferred_field(name):
    key = "_stored_" + name

    def get(self):
        src = self.__dict__.get("_heights_src")
        if src is not None and self.__dict__.get("_heights_gen") != src[0]._state_gen:
            self._materialize_heights()
        return self.__dict__[key]

    def put(self, value):
        self.__dict__[key] = value

    return property(get, put, doc=f"{name} (deferred after Atmosphere_Update; see HEIGHT_FIELDS)")


class Dyn_Data:
    """``Dyn_Data(name, num_fourier, num_spherical, nλ, nθ, nd,
    num_grid_tracters=0, num_spe_tracters=0)`` (Dyn_Data.jl:158-324).  The
    mesh whose stream orders Time_Advance defaults to the last mesh created
    with the same shape (``mesh=`` selects one explicitly).

    The fields of HEIGHT_FIELDS hold, when read, what
    Compute_Pressures_And_Heights! left at the end of the last
    Atmosphere_Update: the step only records that they are due, and the first
    read after it (after an eager step or a graph launch on the mesh) computes
    all seven from grid_ps_c, grid_t_c and grid_geopots and waits for them.
    They therefore match the reference unless those three arrays are changed
    between the step and that first read."""

    def __init__(self, name, num_fourier, num_spherical, nλ, nθ, nd, num_grid_tracters=0,
                 num_spe_tracters=0, mesh=None, device="cuda"):
        self.name = name
        self.num_fourier, self.num_spherical = num_fourier, num_spherical
        self.nλ, self.nθ, self.nd = nλ, nθ, nd
        self.num_grid_tracters, self.num_spe_tracters = num_grid_tracters, num_spe_tracters
        self.mesh = mesh if mesh is not None else current_mesh(num_fourier, nθ, nλ)
        dev = torch.device(device)
        sp = lambda k: torch.zeros((k, num_spherical + 1, num_fourier + 1), dtype=torch.complex128,
                                   device=dev)
        gr = lambda k: torch.zeros((k, nθ, nλ), dtype=torch.float64, device=dev)
        for f in _SPEC_ND:
            setattr(self, f, sp(nd))
        for f in _SPEC_1:
            setattr(self, f, sp(1))
        for f in _GRID_ND:
            setattr(self, f, gr(nd))
        for f in _GRID_1:
            setattr(self, f, gr(1))
        for f in _GRID_HALF:
            setattr(self, f, gr(nd + 1))
        # tracers (Dyn_Data.jl:181-184, 208-211; Julia [.., nd, ntracer] =
        # [ntracer][nd][..] here), allocated even when empty like the reference's
        for s in ("n", "c", "p"):
            setattr(self, f"spe_tracers_{s}", torch.zeros(
                (num_spe_tracters, nd, num_spherical + 1, num_fourier + 1), dtype=torch.complex128,
                device=dev))
            setattr(self, f"grid_tracers_{s}", torch.zeros((num_grid_tracters, nd, nθ, nλ),
                                                           dtype=torch.float64, device=dev))

    def _defer_heights(self, mesh, atmo_data, vert_coord):
        """Record that Compute_Pressures_And_Heights! is due (the caller has
        just enqueued a step on `mesh`)."""
        mesh._state_gen += 1
        self._heights_src = (mesh, atmo_data, vert_coord)

    def _materialize_heights(self):
        mesh, atmo_data, vert_coord = self._heights_src
        self._heights_gen = mesh._state_gen
        from .primitive import Compute_Pressures_And_Heights
        Compute_Pressures_And_Heights(atmo_data, vert_coord, self.grid_ps_c, self.grid_geopots,
                                      self.grid_t_c,
                                      *(self.__dict__["_stored_" + f] for f in HEIGHT_FIELDS))
        if not mesh._capturing:
            mesh.synchronize()


for _f in HEIGHT_FIELDS:
    setattr(Dyn_Data, _f, _deferred_field(_f))


# The prognostic arrays that Time_Advance! advances (Dyn_Data.jl:326-364), in
# its order; a rotated step (Atmosphere_Update(..., rotate=True)) renames
# their p, c, n arrays instead of copying them.
ROTATED_STATE = ["spe_vor", "spe_div", "spe_lnps", "spe_t", "grid_u", "grid_v", "grid_ps",
                 "grid_t"]


def _rotate_state(d, names=None):
    """p <- c, c <- n by renaming the arrays of `names` (default
    ROTATED_STATE): the array that held p becomes the new n (its values are
    stale until the next step writes it).  While a graph is being captured on
    the state's mesh the rotation is logged, so that the graph can check that
    its steps return the names to where they started (Step_Graph)."""
    d._rotated_names = list(names or ROTATED_STATE)
    for f in d._rotated_names:
        p, c, n = getattr(d, f + "_p"), getattr(d, f + "_c"), getattr(d, f + "_n")
        setattr(d, f + "_p", c)
        setattr(d, f + "_c", n)
        setattr(d, f + "_n", p)
    d._rotation = (getattr(d, "_rotation", 0) + 1) % 3
    mesh = d.mesh
    if getattr(mesh, "_capturing", False):
        mesh._rotation_log = getattr(mesh, "_rotation_log", []) + [d]


def _tracer_names(d):
    return [f"{f}_tracers" for f in ("spe", "grid") if getattr(d, f"num_{f}_tracters") > 0]


def Time_Advance(dyn_data):
    """Dyn_Data.jl:326-364: p = c, c = n for the spectral vor, div, lnps, T,
    the grid u, v, ps, T and the tracers (when there are any), as one device
    launch."""
    _advance(dyn_data, [f"spe_{f}" for f in ("vor", "div", "lnps", "t")] +
             [f"grid_{f}" for f in ("u", "v", "ps", "t")] + _tracer_names(dyn_data))


# The prognostic arrays each stepper writes (Time_Advance! copies the others
# too, but in these models they are never written: they stay zero, so their
# copies change nothing and are skipped).
SW_STATE = ["spe_vor", "spe_div", "spe_lnps", "grid_u", "grid_v", "grid_ps"]
BAROTROPIC_STATE = ["spe_vor", "grid_u", "grid_v"]


def Time_Advance_Model(dyn_data, names):
    """Time_Advance! (Dyn_Data.jl:326-364) of the named prognostic arrays and
    the tracers; for the barotropic and shallow-water steps, whose other
    arrays are never written (no reference counterpart)."""
    _advance(dyn_data, list(names) + _tracer_names(dyn_data))


def Time_Advance_Tracers(dyn_data):
    """The tracer part of Time_Advance! alone (Dyn_Data.jl:341-344, 360-363);
    for steps whose other fields are advanced by a fused call.  No launch
    without tracers."""
    names = _tracer_names(dyn_data)
    if names:
        _advance(dyn_data, names)


def _advance(d, names):
    n = len(names)
    ps, cs, ns = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
    counts = (ctypes.c_size_t * n)()
    for i, f in enumerate(names):
        p, c, nx = getattr(d, f + "_p"), getattr(d, f + "_c"), getattr(d, f + "_n")
        ps[i], cs[i], ns[i] = p.data_ptr(), c.data_ptr(), nx.data_ptr()
        counts[i] = c.numel() * c.element_size() // 8
    call("jgcm_time_advance", d.mesh.handle, n, ps, cs, ns, counts)
