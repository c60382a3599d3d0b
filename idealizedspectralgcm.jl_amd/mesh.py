"""Host-side mirror of JGCM's Spectral_Spherical_Mesh API over libjgcm_hip.

Reference: src/Atmos_Spectral/Spectral_Spherical_Mesh.jl (SSM).  Names and
argument order follow the reference; a trailing ``!`` of the Julia name is
dropped (Python identifiers cannot carry it) and the function still mutates
its trailing output arguments in place.  Functions the reference gives a
``cosθ`` argument (Divide_By_Cos!, Multiply_By_Cos!) take the mesh instead.

Fields are device tensors (torch is used only to hold GPU memory) in the
reference's memory layout:

* spectral: ``torch.complex128`` of shape ``(nlev, num_spherical+1, num_fourier+1)``
* grid:     ``torch.float64`` of shape ``(nlev, nθ, nλ)``

with ``nlev == nd`` or ``nlev == 1`` as the reference asserts
(SSM:175,255,370,466,707,724).
"""
import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import call


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class _MeshTable(np.ndarray):
    """Host table that remembers its mesh, so reference-style calls that pass
    a table (e.g. ``Filtered_Leapfrog(..., mesh.laplacian_eig, ...)``) can
    reach the device objects.  The link is a weak reference: the mesh owns its
    tables, and a strong link back would put every mesh in a reference cycle
    that only the cyclic collector frees, in arbitrary order."""
    _mesh_ref = None

    @property
    def mesh(self):
        return None if self._mesh_ref is None else self._mesh_ref()


def _with_mesh(arr, mesh):
    t = np.asarray(arr).view(_MeshTable)
    t._mesh_ref = weakref.ref(mesh)
    return t


_MESHES = []


def current_mesh(num_fourier, nθ, nλ):
    """The most recently created live mesh of this shape."""
    for ref in reversed(_MESHES):
        m = ref()
        if m is not None and m._h is not None and (m.num_fourier, m.nθ, m.nλ) == (num_fourier, nθ, nλ):
            return m
    raise AssertionError(f"no Spectral_Spherical_Mesh with num_fourier={num_fourier}, "
                         f"nθ={nθ}, nλ={nλ}")


class Spectral_Spherical_Mesh:
    """``Spectral_Spherical_Mesh(num_fourier, num_spherical, nθ, nλ, nd, radius)``
    (SSM:69-139).  ``precision`` selects the Legendre-stage arithmetic
    ("fp64", "fp32", "split32"); the reference is FP64."""

    def __init__(self, num_fourier, num_spherical, nθ, nλ, nd, radius, precision="fp64",
                 device="cuda"):
        self.num_fourier, self.num_spherical = int(num_fourier), int(num_spherical)
        self.nθ, self.nλ, self.nd = int(nθ), int(nλ), int(nd)
        self.radius = float(radius)
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        call("jgcm_mesh_create", self.num_fourier, self.num_spherical, self.nθ, self.nλ,
             self.nd, self.radius, ctypes.byref(h))
        self._h = h
        # host-side bookkeeping of deferred diagnostics (Dyn_Data): a counter of
        # state-advancing work enqueued on this mesh (eager steps, graph
        # launches), and whether a graph capture is in progress
        self._state_gen = 0
        self._capturing = False
        self.set_precision(precision)
        N1, S1 = self.num_fourier + 1, self.num_spherical + 1
        shapes = {"sinθ": (self.nθ,), "cosθ": (self.nθ,), "wts": (self.nθ,),
                  "λc": (self.nλ,), "θc": (self.nθ,), "λe": (self.nλ + 1,),
                  "θe": (self.nθ + 1,), "qnm": (N1, S1, self.nθ), "dqnm": (N1, S1, self.nθ),
                  "qwg": (N1, S1, self.nθ), "epsilon": (N1, S1), "laplacian_eig": (N1, S1),
                  "coef_uvm": (N1, S1), "coef_uvc": (N1, S1), "coef_uvp": (N1, S1),
                  "coef_dλ": (N1, S1), "coef_dθm": (N1, S1), "coef_dθp": (N1, S1),
                  "coef_alp_a": (N1, S1, self.nθ), "coef_alp_b": (N1, S1, self.nθ)}
        self._shapes = shapes
        self._tables = {}
        for k in ("sinθ", "cosθ", "wts", "λc", "θc", "λe", "θe", "laplacian_eig", "epsilon"):
            setattr(self, k, _with_mesh(self.table(k), self))
        wn = np.zeros((N1, S1), dtype=np.int64)
        for m in range(N1):
            wn[m, m:] = np.arange(m, S1)
        self.wave_numbers = wn                       # SSM:583-597
        # device copy of cos θ for host-side callers that need it
        self.cosθ_dev = torch.tensor(np.asarray(self.cosθ), dtype=torch.float64, device=self.device)
        self._si_ref = None
        _MESHES.append(weakref.ref(self))

    @property
    def _semi_implicit(self):
        """The most recent Semi_Implicit_Solver built on this mesh (weakly held;
        the default solver of Spectral_Dynamics_Physics)."""
        return None if self._si_ref is None else self._si_ref()

    @_semi_implicit.setter
    def _semi_implicit(self, si):
        self._si_ref = None if si is None else weakref.ref(si)

    # ---- tables (host copies of the reference struct's fields) ----
    def table(self, name):
        if name not in self._tables:
            which = _lib.TABLES[name]
            n = ctypes.c_size_t()
            call("jgcm_mesh_table_size", self._h, which, ctypes.byref(n))
            out = np.empty(n.value, dtype=np.float64)
            call("jgcm_mesh_get_table", self._h, which, out.ctypes.data_as(ctypes.c_void_p), n.value)
            self._tables[name] = out.reshape(self._shapes[name])
        return self._tables[name]

    def __getattr__(self, name):
        if name in ("qnm", "dqnm", "qwg", "coef_uvm", "coef_uvc", "coef_uvp", "coef_dλ",
                    "coef_dθm", "coef_dθp", "coef_alp_a", "coef_alp_b"):
            return self.table(name)
        raise AttributeError(name)

    # ---- handle management ----
    @property
    def handle(self):
        return self._h

    def set_precision(self, precision):
        call("jgcm_mesh_set_precision", self._h, _lib.PREC[precision])
        self.precision = precision

    def stream(self):
        return _lib.lib().jgcm_mesh_stream(self._h)

    def synchronize(self):
        call("jgcm_mesh_synchronize", self._h)

    def reserve(self, nfields):
        """Grow the mesh scratch to `nfields` level-fields (needed before graph capture)."""
        call("jgcm_mesh_reserve", self._h, int(nfields))

    def lifetime(self):
        """(references, live graphs) of the library's mesh object."""
        r, g = ctypes.c_int(), ctypes.c_int()
        call("jgcm_mesh_lifetime", self._h, ctypes.byref(r), ctypes.byref(g))
        return r.value, g.value

    def close(self):
        """Drops this object's reference; leapfrogs, vertical coordinates,
        graphs and meshes borrowing the stream keep the library object alive
        until they are destroyed."""
        if getattr(self, "_h", None) is not None and self._h.value:
            call("jgcm_mesh_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- allocation helpers ----
    def spectral_zeros(self, nlev=None):
        nlev = self.nd if nlev is None else nlev
        return torch.zeros((nlev, self.num_spherical + 1, self.num_fourier + 1),
                           dtype=torch.complex128, device=self.device)

    def grid_zeros(self, nlev=None):
        nlev = self.nd if nlev is None else nlev
        return torch.zeros((nlev, self.nθ, self.nλ), dtype=torch.float64, device=self.device)

    # ---- argument checks mirroring the reference asserts ----
    def _nlev(self, t, kind, batch=False):
        assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous(), \
            f"{kind} field must be a contiguous device tensor"
        if kind == "spec":
            assert t.dtype == torch.complex128 and t.dim() == 3 and \
                tuple(t.shape[1:]) == (self.num_spherical + 1, self.num_fourier + 1), \
                f"spectral field shape {tuple(t.shape)}"
        elif kind == "grid":
            assert t.dtype == torch.float64 and t.dim() == 3 and \
                tuple(t.shape[1:]) == (self.nθ, self.nλ), f"grid field shape {tuple(t.shape)}"
        elif kind == "four":
            assert t.dtype == torch.complex128 and t.dim() == 3 and \
                tuple(t.shape[1:]) == (self.nθ, self.nλ), f"fourier field shape {tuple(t.shape)}"
        nlev = t.shape[0]
        if not batch:
            assert nlev == self.nd or nlev == 1, f"level count {nlev} is neither nd={self.nd} nor 1"
        return nlev


def _same(*ts):
    n = ts[0].shape[0]
    assert all(t.shape[0] == n for t in ts), "level counts differ"
    return n


# --------------------------------------------------------------- transforms
def Trans_Spherical_To_Grid(mesh, snm, pfield, batch=False):
    """SSM:143-217 (``batch=True`` lifts the nd-or-1 level restriction)."""
    n = mesh._nlev(snm, "spec", batch)
    mesh._nlev(pfield, "grid", batch)
    _same(snm, pfield)
    call("jgcm_trans_spherical_to_grid", mesh.handle, _ptr(snm), _ptr(pfield), n)


def Trans_Grid_To_Spherical(mesh, pfield, snm, batch=False):
    """SSM:221-293."""
    n = mesh._nlev(pfield, "grid", batch)
    mesh._nlev(snm, "spec", batch)
    _same(snm, pfield)
    call("jgcm_trans_grid_to_spherical", mesh.handle, _ptr(pfield), _ptr(snm), n)


BATCH_DIVCOS = 1
BATCH_VORDIV = 2
INV_GRAD_LAMBDA = 2
INV_GRAD_THETA = 4
INV_UCOS = 6
INV_VCOS = 8


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[None if t is None else t.data_ptr() for t in ts])


def _batch_args(ts_a, ts_b, flags):
    n = len(ts_a)
    assert 1 <= n <= 12 and len(ts_b) == n, "batched transforms take 1..12 arrays"
    nl = (ctypes.c_int * n)(*[t.shape[0] for t in ts_a])
    fl = (ctypes.c_int * n)(*(flags or [0] * n))
    return n, _ptrs(ts_a), _ptrs(ts_b), nl, fl


def Trans_Spherical_To_Grid_Batch(mesh, snms, pfields, flags=None, snms_b=None):
    """Trans_Spherical_To_Grid! of several arrays in one Legendre launch and
    one FFT launch (no reference counterpart).  flags[k] may combine
    BATCH_DIVCOS (divide array k's rows by cos θ) with one spectral operator
    applied on the way in: INV_GRAD_LAMBDA / INV_GRAD_THETA (cos-weighted
    gradient, SSM:655-692) or INV_UCOS / INV_VCOS (u cos θ, v cos θ from
    vor = snms[k] and div = snms_b[k], SSM:517-580) -- DIVCOS with those is
    Compute_Gradients! and UV_Grid_From_Vor_Div!."""
    for k, (s, g) in enumerate(zip(snms, pfields)):
        mesh._nlev(s, "spec", True)
        mesh._nlev(g, "grid", True)
        _same(s, g)
        if snms_b is not None and snms_b[k] is not None:
            mesh._nlev(snms_b[k], "spec", True)
            _same(s, snms_b[k])
    n, a, b, nl, fl = _batch_args(snms, pfields, flags)
    sb = _ptrs(snms_b) if snms_b is not None else None
    call("jgcm_trans_spherical_to_grid_batch", mesh.handle, n, a, sb, b, nl, fl)


def Trans_Grid_To_Spherical_Batch(mesh, pfields, snms, flags=None):
    """Trans_Grid_To_Spherical! of several arrays in one FFT launch and one
    Legendre launch (no reference counterpart).  flags[k] = BATCH_VORDIV
    makes arrays k, k+1 a (u, v) pair whose outputs are (vor, div), as
    Vor_Div_From_Grid_UV!."""
    for s, g in zip(snms, pfields):
        mesh._nlev(s, "spec", True)
        mesh._nlev(g, "grid", True)
        _same(s, g)
    call("jgcm_trans_grid_to_spherical_batch", mesh.handle, *_batch_args(pfields, snms, flags))


def Trans_Grid_To_Fourier(mesh, pfield, fourier_g):
    """SSM:295-330."""
    n = mesh._nlev(pfield, "grid", True)
    mesh._nlev(fourier_g, "four", True)
    _same(pfield, fourier_g)
    call("jgcm_trans_grid_to_fourier", mesh.handle, _ptr(pfield), _ptr(fourier_g), n)


def Divide_By_Cos(mesh, grid_d):
    """SSM:333-338."""
    n = mesh._nlev(grid_d, "grid", True)
    call("jgcm_divide_by_cos", mesh.handle, _ptr(grid_d), n)


def Multiply_By_Cos(mesh, grid_d, grid_d_cos):
    """SSM:341-346."""
    n = mesh._nlev(grid_d, "grid", True)
    mesh._nlev(grid_d_cos, "grid", True)
    _same(grid_d, grid_d_cos)
    call("jgcm_multiply_by_cos", mesh.handle, _ptr(grid_d), _ptr(grid_d_cos), n)


def Vor_Div_From_Grid_UV(mesh, grid_u, grid_v, vor, div, batch=False):
    """SSM:349-393."""
    n = mesh._nlev(grid_u, "grid", batch)
    for t, k in ((grid_v, "grid"), (vor, "spec"), (div, "spec")):
        mesh._nlev(t, k, batch)
    _same(grid_u, grid_v, vor, div)
    call("jgcm_vor_div_from_grid_uv", mesh.handle, _ptr(grid_u), _ptr(grid_v), _ptr(vor),
         _ptr(div), n)


def Compute_Alpha_Operator(mesh, fourier_a, fourier_b, isign, alpha):
    """SSM:424-459 (fourier_a/b as produced by Trans_Grid_To_Fourier)."""
    n = mesh._nlev(fourier_a, "four", True)
    mesh._nlev(fourier_b, "four", True)
    mesh._nlev(alpha, "spec", True)
    _same(fourier_a, fourier_b, alpha)
    call("jgcm_compute_alpha_operator", mesh.handle, _ptr(fourier_a), _ptr(fourier_b),
         float(isign), _ptr(alpha), n)


def UV_Grid_From_Vor_Div(mesh, vor, div, grid_u, grid_v, batch=False):
    """SSM:462-482."""
    n = mesh._nlev(vor, "spec", batch)
    for t, k in ((div, "spec"), (grid_u, "grid"), (grid_v, "grid")):
        mesh._nlev(t, k, batch)
    _same(vor, div, grid_u, grid_v)
    call("jgcm_uv_grid_from_vor_div", mesh.handle, _ptr(vor), _ptr(div), _ptr(grid_u),
         _ptr(grid_v), n)


def Compute_Ucos_Vcos_From_Vor_Div(mesh, vor, div, spherical_ucos, spherical_vcos):
    """SSM:517-580."""
    n = mesh._nlev(vor, "spec", True)
    for t in (div, spherical_ucos, spherical_vcos):
        mesh._nlev(t, "spec", True)
    _same(vor, div, spherical_ucos, spherical_vcos)
    call("jgcm_compute_ucos_vcos_from_vor_div", mesh.handle, _ptr(vor), _ptr(div),
         _ptr(spherical_ucos), _ptr(spherical_vcos), n)


def Apply_Laplacian(mesh, spherical_u):
    """SSM:614-621."""
    n = mesh._nlev(spherical_u, "spec", True)
    call("jgcm_apply_laplacian", mesh.handle, _ptr(spherical_u), n)


def Compute_Gradient_Cos(mesh, spe_hs, spe_cos_dλ_hs, spe_cos_dθ_hs):
    """SSM:655-692."""
    n = mesh._nlev(spe_hs, "spec", True)
    mesh._nlev(spe_cos_dλ_hs, "spec", True)
    mesh._nlev(spe_cos_dθ_hs, "spec", True)
    _same(spe_hs, spe_cos_dλ_hs, spe_cos_dθ_hs)
    call("jgcm_compute_gradient_cos", mesh.handle, _ptr(spe_hs), _ptr(spe_cos_dλ_hs),
         _ptr(spe_cos_dθ_hs), n)


def Compute_Gradients(mesh, spe_hs, grid_dλ_hs, grid_dθ_hs):
    """SSM:721-741."""
    n = mesh._nlev(spe_hs, "spec")
    mesh._nlev(grid_dλ_hs, "grid")
    mesh._nlev(grid_dθ_hs, "grid")
    _same(spe_hs, grid_dλ_hs, grid_dθ_hs)
    call("jgcm_compute_gradients", mesh.handle, _ptr(spe_hs), _ptr(grid_dλ_hs),
         _ptr(grid_dθ_hs), n)


def Add_Horizontal_Advection(mesh, spe_hs, grid_u, grid_v, grid_δhs):
    """SSM:694-718."""
    n = mesh._nlev(spe_hs, "spec")
    for t in (grid_u, grid_v, grid_δhs):
        mesh._nlev(t, "grid")
    _same(spe_hs, grid_u, grid_v, grid_δhs)
    call("jgcm_add_horizontal_advection", mesh.handle, _ptr(spe_hs), _ptr(grid_u), _ptr(grid_v),
         _ptr(grid_δhs), n)


def Area_Weighted_Global_Mean(mesh, grid_datas):
    """SSM:746-760 (level 1)."""
    mesh._nlev(grid_datas, "grid", True)
    out = ctypes.c_double()
    call("jgcm_area_weighted_global_mean", mesh.handle, _ptr(grid_datas), ctypes.byref(out))
    return out.value
