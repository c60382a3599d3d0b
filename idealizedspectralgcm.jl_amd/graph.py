"""HIP-graph capture of whole steps (no reference counterpart: the reference
runs on the CPU).  Any sequence of library calls on one mesh can be captured
and replayed; the captured kernels keep the argument values of capture time,
so a leapfrog step is captured once for the first (init) step and once for
the later steps, whose Δt no longer changes.
"""
import ctypes
import gc

from ._lib import call


def _unrotate(log):
    """Undo the renamings a capture logged (nothing of it ran on the device)."""
    from .dyn_data import _rotate_state
    for d in {id(d): d for d in log}.values():
        k = sum(1 for x in log if x is d) % 3
        for _ in range((3 - k) % 3):
            _rotate_state(d, d._rotated_names)


class Step_Graph:
    def __init__(self, mesh, fn, nfields=None, repeat=1):
        """Capture ``fn()`` (which must only enqueue library calls on `mesh`),
        ``repeat`` times in a row: one launch then advances the state by
        ``repeat`` steps, and the launch overhead between graphs is paid once
        per ``repeat`` steps.  ``fn`` must enqueue the same work each time
        (a later leapfrog step, not the init step).

        The cyclic garbage collector is flushed and paused for the capture: a
        collection inside it could finalize an unrelated mesh or solver, whose
        destroy call (stream synchronisation, hipFree) is not allowed while
        this thread captures and would invalidate the capture."""
        self.mesh = mesh
        if nfields:
            mesh.reserve(nfields)
        gc.collect()
        was_enabled = gc.isenabled()
        gc.disable()
        mesh._rotation_log = []
        try:
            mesh.synchronize()
            call("jgcm_graph_begin", mesh.handle)
            mesh._capturing = True
            try:
                for _ in range(int(repeat)):
                    fn()
            except Exception:
                g = ctypes.c_void_p()
                try:
                    call("jgcm_graph_end", mesh.handle, ctypes.byref(g))
                    call("jgcm_graph_destroy", g)
                except Exception:
                    pass
                mesh._capturing = False
                _unrotate(mesh.__dict__.pop("_rotation_log", []))
                raise
            g = ctypes.c_void_p()
            call("jgcm_graph_end", mesh.handle, ctypes.byref(g))
        finally:
            mesh._capturing = False
            if was_enabled:
                gc.enable()
        self._g = g
        self.repeat = int(repeat)
        # Rotated steps (Atmosphere_Update(..., rotate=True)) rename their
        # state arrays, and the captured kernels keep the arrays of capture
        # time: a graph can be replayed only if its steps bring every renamed
        # state back to the names it started with (a multiple of 3 rotations),
        # and only while the state still has those names.
        log = mesh.__dict__.pop("_rotation_log", [])
        self._states = []
        if any(sum(1 for x in log if x is d) % 3 for d in log):
            _unrotate(log)
            self.close()
            raise AssertionError("the captured steps rotate a state a number of times that is "
                                 "not a multiple of 3; rotated steps need graphs of a multiple "
                                 "of 3 steps")
        for d in {id(d): d for d in log}.values():
            self._states.append((d, d._rotation))

    def launch(self):
        for d, phase in self._states:
            assert d._rotation == phase, \
                "the state was rotated since the capture (an eager rotated step in between)"
        call("jgcm_graph_launch", self._g)
        self.mesh._state_gen += 1   # the state advanced: deferred diagnostics are stale

    def close(self):
        if getattr(self, "_g", None) is not None and self._g.value:
            call("jgcm_graph_destroy", self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
