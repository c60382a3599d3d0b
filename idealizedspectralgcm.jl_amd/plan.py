"""Transform plans: the batched transforms of several meshes in one launch per
stage (no reference counterpart: the reference steps one model at a time).

A plan is built once for fixed arrays (like an FFT plan) and executed as often
as needed, eagerly or inside a captured graph.  Its Legendre stage takes every
tile of the one-mesh batched launches -- (mesh, m, parity, row or latitude
group, field chunk) -- from one cost-sorted work list in ONE launch, dealt
longest-first to persistent workgroups; the FFT stage runs per mesh.  Results
equal `Trans_Spherical_To_Grid_Batch` / `Trans_Grid_To_Spherical_Batch` on each
mesh bit for bit (include/jgcm.h, jgcm_plan_*).
"""
import ctypes

import numpy as np

from ._lib import call, lib
from .mesh import _ptrs, _same

INVERSE, FORWARD = "inverse", "forward"


def _int_array(values):
    return (ctypes.c_int * len(values))(*values)


class Transform_Plan:
    """``direction``: ``"inverse"`` (spectral -> grid) with batches
    ``(mesh, snms, pfields, flags, snms_b)`` or ``"forward"`` (grid ->
    spectral) with batches ``(mesh, pfields, snms, flags)``, one per mesh;
    the arguments per mesh are those of the one-mesh batched calls.  The
    meshes must share one stream and one Legendre precision.
    ``workgroups``: > 0 persistent workgroups, < 0 as many as the device
    holds at once, 0 one workgroup per work item."""

    def __init__(self, direction, batches, workgroups=-1):
        assert direction in (INVERSE, FORWARD), direction
        assert 1 <= len(batches) <= 8, "a plan takes 1..8 meshes"
        self.direction = direction
        self.meshes = [b[0] for b in batches]
        self._keep = []        # every array the plan addresses stays alive with it
        nseg, nlev, flags, a_arr, b_arr, c_arr = [], [], [], [], [], []
        for b in batches:
            if direction == INVERSE:
                mesh, snms, pfields, fl, snms_b = (tuple(b) + (None, None))[:5]
                first, second = snms, pfields
            else:
                mesh, pfields, snms, fl = (tuple(b) + (None,))[:4]
                first, second, snms_b = pfields, snms, None
            n = len(first)
            assert 1 <= n <= 12 and len(second) == n, "a batch takes 1..12 arrays"
            for k, (s, g) in enumerate(zip(snms, pfields)):
                mesh._nlev(s, "spec", True)
                mesh._nlev(g, "grid", True)
                _same(s, g)
                if snms_b is not None and snms_b[k] is not None:
                    mesh._nlev(snms_b[k], "spec", True)
                    _same(s, snms_b[k])
            nseg.append(n)
            nlev.append(_int_array([t.shape[0] for t in first]))
            flags.append(_int_array(list(fl) if fl else [0] * n))
            a_arr.append(_ptrs(first))
            b_arr.append(_ptrs(second))
            c_arr.append(_ptrs(snms_b) if snms_b is not None else None)
            self._keep += [t for t in list(first) + list(second) + list(snms_b or []) if t is not None]
        nm = len(batches)
        P = ctypes.c_void_p
        meshes = (P * nm)(*[m.handle for m in self.meshes])
        ptr_of = lambda arrs: (P * nm)(*[None if a is None else ctypes.addressof(a) for a in arrs])
        h = P()
        if direction == INVERSE:
            call("jgcm_plan_spherical_to_grid_create", nm, meshes, _int_array(nseg), ptr_of(a_arr),
                 ptr_of(c_arr), ptr_of(b_arr), ptr_of(nlev), ptr_of(flags), int(workgroups),
                 ctypes.byref(h))
        else:
            call("jgcm_plan_grid_to_spherical_create", nm, meshes, _int_array(nseg), ptr_of(a_arr),
                 ptr_of(b_arr), ptr_of(nlev), ptr_of(flags), int(workgroups), ctypes.byref(h))
        self.handle = h
        self._lib = lib()

    def execute(self):
        call("jgcm_plan_execute", self.handle)

    def info(self):
        """(work items, workgroups) of the Legendre launch."""
        n, w = ctypes.c_int(), ctypes.c_int()
        call("jgcm_plan_info", self.handle, ctypes.byref(n), ctypes.byref(w))
        return n.value, w.value

    def work_list(self):
        """(items [n, 4] = (mesh, 2m + parity, y, z) in launch order,
        wg_start [workgroups + 1])."""
        n, w = self.info()
        items = np.zeros((max(n, 1), 4), np.int32)
        starts = np.zeros(w + 1, np.int32)
        call("jgcm_plan_work_list", self.handle, items.ctypes.data_as(ctypes.c_void_p),
             starts.ctypes.data_as(ctypes.c_void_p))
        return items[:n], starts

    def close(self):
        if getattr(self, "handle", None):
            self._lib.jgcm_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def LPT_Schedule(cost, workgroups):
    """The plans' scheduler (host only): items in decreasing cost to the
    least-loaded workgroup.  Returns (order, wg_start)."""
    cost = np.ascontiguousarray(cost, dtype=np.float64)
    n = cost.size
    order = np.zeros(max(n, 1), np.int32)
    starts = np.zeros(max(n, int(workgroups), 0) + 1, np.int32)
    used = ctypes.c_int()
    call("jgcm_lpt_schedule", n, cost.ctypes.data_as(ctypes.c_void_p), int(workgroups),
         order.ctypes.data_as(ctypes.c_void_p), starts.ctypes.data_as(ctypes.c_void_p),
         ctypes.byref(used))
    return order[:n], starts[:used.value + 1]
