"""Filtered_Leapfrog mirror over the C ABI.

Reference: src/Atmos_Spectral/Time_Integrator.jl (TI).  The integrator's
damping table and leapfrog/Robert-Asselin update run on the device; the
``init_step`` flag and Δt live in the C object, ``time`` is a host counter the
drivers advance themselves (``integrator.time += Δt``, HS.jl:75).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call
from .mesh import _ptr


class Filtered_Leapfrog:
    """``Filtered_Leapfrog(robert_coef, damping_order, damping_coef, eigen,
    implicit_coef, Δt, init_step, start_time, end_time)`` (TI:20-38).
    ``eigen`` must be a mesh's ``laplacian_eig`` (it carries the mesh)."""

    def __init__(self, robert_coef, damping_order, damping_coef, eigen, implicit_coef, Δt,
                 init_step, start_time, end_time):
        mesh = getattr(eigen, "mesh", None)
        assert mesh is not None, "eigen must be Spectral_Spherical_Mesh.laplacian_eig"
        assert damping_order % 2 == 0                                      # TI:24
        self.mesh = mesh
        h = ctypes.c_void_p()
        call("jgcm_leapfrog_create", mesh.handle, float(robert_coef), int(damping_order),
             float(damping_coef), float(implicit_coef), int(Δt), 1 if init_step else 0,
             int(start_time), int(end_time), ctypes.byref(h))
        self._h = h
        self.robert_coef = float(robert_coef)
        self.damping_order = int(damping_order)
        self.damping_coef = float(damping_coef)
        self.laplacian_eigen = eigen
        self.implicit_coef = float(implicit_coef)
        self.Δt = int(Δt)
        self.time = int(start_time)
        self.start_time, self.end_time = int(start_time), int(end_time)
        d = np.empty(eigen.size, dtype=np.float64)
        call("jgcm_leapfrog_damping", h, d.ctypes.data_as(ctypes.c_void_p), d.size)
        self.damping = d.reshape(eigen.shape)

    @property
    def handle(self):
        return self._h

    @property
    def init_step(self):
        flag = ctypes.c_int()
        call("jgcm_leapfrog_state", self._h, ctypes.byref(flag), None, None)
        return bool(flag.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("jgcm_leapfrog_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def Update_Init_Step(integrator):
    """TI:40-43 (the semi-implicit overload lives in semi_implicit.py)."""
    if hasattr(integrator, "semi_implicit_update"):
        return integrator.semi_implicit_update()
    call("jgcm_leapfrog_update_init_step", integrator.handle)


def Get_Δt(integrator):
    """TI:46-50."""
    return integrator.Δt if integrator.init_step else 2 * integrator.Δt


def Get_ξ(integrator):
    """TI:52-56."""
    return (integrator.Δt * integrator.implicit_coef if integrator.init_step
            else 2 * integrator.Δt * integrator.implicit_coef)


def _spec_n(integrator, *ts):
    mesh = integrator.mesh
    n = mesh._nlev(ts[0], "spec", True)
    for t in ts[1:]:
        assert mesh._nlev(t, "spec", True) == n, "level counts differ"
    return n


def Compute_Spectral_Damping(integrator, Qc, Qp, δQ):
    """TI:58-81."""
    n = _spec_n(integrator, Qc, Qp, δQ)
    call("jgcm_compute_spectral_damping", integrator.handle, _ptr(Qc), _ptr(Qp), _ptr(δQ), n)


def Filtered_Leapfrog_step(integrator, δQ, Qp, Qc, Qn):
    """``Filtered_Leapfrog!`` TI:86-101 (the Julia name is also the struct's;
    the Python mirror names the function ``Filtered_Leapfrog_step``)."""
    n = _spec_n(integrator, δQ, Qp, Qc, Qn)
    call("jgcm_filtered_leapfrog", integrator.handle, _ptr(δQ), _ptr(Qp), _ptr(Qc), _ptr(Qn), n)


def Damped_Leapfrog(integrator, δQ, Qp, Qc, Qn):
    """Compute_Spectral_Damping! then Filtered_Leapfrog! in one kernel (same results)."""
    n = _spec_n(integrator, δQ, Qp, Qc, Qn)
    call("jgcm_damped_leapfrog", integrator.handle, _ptr(δQ), _ptr(Qp), _ptr(Qc), _ptr(Qn), n)


def Damped_Leapfrog_Batch(integrator, fields):
    """Compute_Spectral_Damping! and/or Filtered_Leapfrog! of up to 8 fields in
    one launch (no reference counterpart).  ``fields`` holds tuples
    ``(δQ, Qp, Qc, Qn, damp)``: ``damp`` True applies the damping before the
    leapfrog (Damped_Leapfrog), False only the leapfrog."""
    n = len(fields)
    assert 1 <= n <= 8, "1..8 fields per batch"
    P = ctypes.c_void_p * n
    dq, qp, qc, qn = P(), P(), P(), P()
    nl, mode = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    for i, (δQ, Qp, Qc, Qn, damp) in enumerate(fields):
        nl[i] = _spec_n(integrator, δQ, Qp, Qc, Qn)
        dq[i], qp[i], qc[i], qn[i] = δQ.data_ptr(), Qp.data_ptr(), Qc.data_ptr(), Qn.data_ptr()
        mode[i] = 3 if damp else 2
    call("jgcm_damped_leapfrog_batch", integrator.handle, n, dq, qp, qc, qn, nl, mode)
