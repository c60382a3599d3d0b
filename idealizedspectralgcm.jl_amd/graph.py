"""HIP-graph capture of whole steps (no reference counterpart: the reference
runs on the CPU).  Any sequence of library calls on one mesh can be captured
and replayed; the captured kernels keep the argument values of capture time,
so a leapfrog step is captured once for the first (init) step and once for
the later steps, whose Δt no longer changes.
"""
import ctypes
import gc

from ._lib import call


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
                raise
            g = ctypes.c_void_p()
            call("jgcm_graph_end", mesh.handle, ctypes.byref(g))
        finally:
            mesh._capturing = False
            if was_enabled:
                gc.enable()
        self._g = g
        self.repeat = int(repeat)

    def launch(self):
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
