"""Reverse-mode derivatives: the transform pair and a barotropic run.

The transforms are linear, so their vector-Jacobian products are the adjoint
transforms of the C ABI (weighted-transpose Legendre + conjugate FFT).  They
are exposed as ``torch.autograd.Function`` subclasses, the Python counterpart
of a ChainRulesCore ``rrule`` on the Julia side; gradients are with respect to
the real grid values and, for spectral arrays, with respect to the real and
imaginary parts (torch's convention for complex tensors, <a, b> = Re(conj(a) b)).

``Barotropic_Adjoint`` and ``Shallow_Water_Adjoint`` differentiate a whole
run of ``Barotropic_Dynamics!`` (Barotropic_Dynamics.jl:2-46) or of
``Shallow_Water_Physics!`` + ``Shallow_Water_Dynamics!``
(Shallow_Water_Dynamics.jl:7-179), with ``Update_Init_Step!`` after the first
step as in exp/Barotropic and exp/Shallow_Water, with respect to the initial
spectral state: a forward pass that keeps the state every
``checkpoint_every`` steps, then a reverse sweep that recomputes each segment
forward (keeping its grid fields) and runs the C step adjoint
(``jgcm_barotropic_step_adjoint``, ``jgcm_sw_step_adjoint``) backwards through
it.  Device memory is O(nsteps / checkpoint_every + checkpoint_every) states.
"""
import torch

from ._lib import call
from .mesh import _ptr


def trans_grid_to_spherical_adjoint(mesh, snm, pfield):
    n = mesh._nlev(snm, "spec", True)
    mesh._nlev(pfield, "grid", True)
    call("jgcm_trans_grid_to_spherical_adjoint", mesh.handle, _ptr(snm), _ptr(pfield), n)


def trans_spherical_to_grid_adjoint(mesh, pfield, snm):
    n = mesh._nlev(pfield, "grid", True)
    mesh._nlev(snm, "spec", True)
    call("jgcm_trans_spherical_to_grid_adjoint", mesh.handle, _ptr(pfield), _ptr(snm), n)


class GridToSpherical(torch.autograd.Function):
    """y = Trans_Grid_To_Spherical!(mesh, x) as a differentiable function."""

    @staticmethod
    def forward(ctx, mesh, grid):
        ctx.mesh = mesh
        out = mesh.spectral_zeros(grid.shape[0])
        call("jgcm_trans_grid_to_spherical", mesh.handle, _ptr(grid.contiguous()), _ptr(out),
             grid.shape[0])
        mesh.synchronize()
        return out

    @staticmethod
    def backward(ctx, g):
        mesh = ctx.mesh
        g = g.contiguous()
        out = mesh.grid_zeros(g.shape[0])
        trans_grid_to_spherical_adjoint(mesh, g, out)
        mesh.synchronize()
        return None, out


class SphericalToGrid(torch.autograd.Function):
    """y = Trans_Spherical_To_Grid!(mesh, s) as a differentiable function."""

    @staticmethod
    def forward(ctx, mesh, snm):
        ctx.mesh = mesh
        out = mesh.grid_zeros(snm.shape[0])
        call("jgcm_trans_spherical_to_grid", mesh.handle, _ptr(snm.contiguous()), _ptr(out),
             snm.shape[0])
        mesh.synchronize()
        return out

    @staticmethod
    def backward(ctx, g):
        mesh = ctx.mesh
        g = g.contiguous()
        out = mesh.spectral_zeros(g.shape[0])
        trans_spherical_to_grid_adjoint(mesh, g, out)
        mesh.synchronize()
        return None, out


# ---- operator adjoints of the barotropic step --------------------------------

def vor_div_from_grid_uv_adjoint(mesh, vor_bar, div_bar, u_bar, v_bar):
    """Transpose of Vor_Div_From_Grid_UV! (SSM:349-393); div_bar may be None."""
    n = mesh._nlev(vor_bar, "spec", True)
    if div_bar is not None:
        assert mesh._nlev(div_bar, "spec", True) == n
    assert mesh._nlev(u_bar, "grid", True) == n and mesh._nlev(v_bar, "grid", True) == n
    call("jgcm_vor_div_from_grid_uv_adjoint", mesh.handle, _ptr(vor_bar),
         None if div_bar is None else _ptr(div_bar), _ptr(u_bar), _ptr(v_bar), n)


def ucos_vcos_adjoint(mesh, ucos_bar, vcos_bar, vor_bar, div_bar=None):
    """Transpose of Compute_Ucos_Vcos_From_Vor_Div! (SSM:517-580)."""
    n = mesh._nlev(ucos_bar, "spec", True)
    for t in (vcos_bar, vor_bar) + (() if div_bar is None else (div_bar,)):
        assert mesh._nlev(t, "spec", True) == n
    call("jgcm_ucos_vcos_adjoint", mesh.handle, _ptr(ucos_bar), _ptr(vcos_bar), _ptr(vor_bar),
         None if div_bar is None else _ptr(div_bar), n)


def uv_grid_from_vor_div_adjoint(mesh, u_bar, v_bar, vor_bar, div_bar=None):
    """Transpose of UV_Grid_From_Vor_Div! (SSM:462-482)."""
    n = mesh._nlev(u_bar, "grid", True)
    assert mesh._nlev(v_bar, "grid", True) == n and mesh._nlev(vor_bar, "spec", True) == n
    if div_bar is not None:
        assert mesh._nlev(div_bar, "spec", True) == n
    call("jgcm_uv_grid_from_vor_div_adjoint", mesh.handle, _ptr(u_bar), _ptr(v_bar), _ptr(vor_bar),
         None if div_bar is None else _ptr(div_bar), n)


def vorticity_flux_adjoint(mesh, grid_vor, grid_u, grid_v, omega, du_bar, dv_bar, vor_bar, u_bar,
                           v_bar):
    """Transpose of δu = (ζ+f) v, δv = -(ζ+f) u (Barotropic_Dynamics.jl:26-29) at (ζ, u, v)."""
    n = mesh._nlev(grid_vor, "grid", True)
    for t in (grid_u, grid_v, du_bar, dv_bar, vor_bar, u_bar, v_bar):
        assert mesh._nlev(t, "grid", True) == n
    call("jgcm_vorticity_flux_adjoint", mesh.handle, _ptr(grid_vor), _ptr(grid_u), _ptr(grid_v),
         float(omega), _ptr(du_bar), _ptr(dv_bar), _ptr(vor_bar), _ptr(u_bar), _ptr(v_bar), n)


def damped_leapfrog_adjoint(integrator, init_step, p_bar_next, c_bar_next, p_bar, c_bar, g_bar):
    """Transpose of Compute_Spectral_Damping! + Filtered_Leapfrog! + Time_Advance!
    (Time_Integrator.jl:58-101): (p_bar', c_bar') -> (p_bar, c_bar, g_bar)."""
    mesh = integrator.mesh
    n = mesh._nlev(p_bar_next, "spec", True)
    for t in (c_bar_next, p_bar, c_bar, g_bar):
        assert mesh._nlev(t, "spec", True) == n
    call("jgcm_damped_leapfrog_adjoint", integrator.handle, 1 if init_step else 0, _ptr(p_bar_next),
         _ptr(c_bar_next), _ptr(p_bar), _ptr(c_bar), _ptr(g_bar), n)


def sw_step_adjoint(integrator, init_step, grids, spe_h, omega, kappa_m, kappa_t, h_0, bars):
    """One reverse step of Shallow_Water_Physics! + Shallow_Water_Dynamics!
    (Shallow_Water_Dynamics.jl:7-21, :88-179), in place on
    bars = (vor_p, vor_c, div_p, div_c, h_p, h_c) cotangents; grids =
    (vor, div, h, u, v) of the step's starting state."""
    mesh = integrator.mesh
    n = mesh._nlev(spe_h, "spec", True)
    assert len(grids) == 5 and len(bars) == 6
    for t in grids:
        assert mesh._nlev(t, "grid", True) == n
    for t in bars:
        assert mesh._nlev(t, "spec", True) == n
    call("jgcm_sw_step_adjoint", integrator.handle, 1 if init_step else 0,
         *[_ptr(t) for t in grids], _ptr(spe_h), float(omega), float(kappa_m), float(kappa_t),
         float(h_0), *[_ptr(t) for t in bars], n)


def barotropic_step_adjoint(integrator, init_step, grid_vor, grid_u, grid_v, omega, p_bar, c_bar):
    """One reverse step of Barotropic_Dynamics!, in place on (p_bar, c_bar)."""
    mesh = integrator.mesh
    n = mesh._nlev(p_bar, "spec", True)
    assert mesh._nlev(c_bar, "spec", True) == n
    for t in (grid_vor, grid_u, grid_v):
        assert mesh._nlev(t, "grid", True) == n
    call("jgcm_barotropic_step_adjoint", integrator.handle, 1 if init_step else 0, _ptr(grid_vor),
         _ptr(grid_u), _ptr(grid_v), float(omega), _ptr(p_bar), _ptr(c_bar), n)


class _Run_Adjoint:
    """Checkpointed reverse sweep over a run of a one-level stepper whose
    state is a tuple of spectral fields.  Subclasses name the state fields
    (``STATE``, pairs of Dyn_Data attribute names in (p, c) order), rebuild the
    grid fields from the spectral state (``_load_grids``), run one step
    (``_step``), snapshot what the step adjoint needs (``_snapshot``) and call
    the C step adjoint (``_step_adjoint``)."""

    STATE = ()

    def __init__(self, mesh, atmo_data, integrator, checkpoint_every):
        from .dyn_data import Dyn_Data
        from .time_integrator import Filtered_Leapfrog
        assert checkpoint_every >= 1
        self.mesh, self.atmo = mesh, atmo_data
        self.every = int(checkpoint_every)
        self.starts_init = bool(integrator.init_step)

        def clone(init):
            return Filtered_Leapfrog(integrator.robert_coef, integrator.damping_order,
                                     integrator.damping_coef, integrator.laplacian_eigen,
                                     integrator.implicit_coef, integrator.Δt, init,
                                     integrator.start_time, integrator.end_time)
        self._first = clone(self.starts_init)
        self._later = clone(False)
        self._d = Dyn_Data("adjoint", mesh.num_fourier, mesh.num_spherical, mesh.nλ, mesh.nθ, 1,
                           mesh=mesh)
        self._checkpoints = []
        self.nsteps = 0

    def _integrator(self, k):
        return self._first if k == 0 else self._later

    def _load(self, state):
        for name, t in zip(self.STATE, state):
            getattr(self._d, name).copy_(t)
        self._load_grids()

    def _state(self):
        return tuple(getattr(self._d, name).clone() for name in self.STATE)

    def _run(self, state, nsteps):
        assert nsteps >= 1
        assert len(state) == len(self.STATE)
        for t in state:
            self.mesh._nlev(t, "spec", True)
            assert t.shape[0] == 1, "the state has one level"
        self._checkpoints = []
        self.nsteps = int(nsteps)
        self._load(state)
        for k in range(self.nsteps):
            if k % self.every == 0:
                self._checkpoints.append(self._state())
            self._step(k)
        return self._state()

    def _reverse(self, bars):
        assert self._checkpoints, "forward() first"
        assert len(bars) == len(self.STATE)
        bars = tuple(b.clone() for b in bars)
        for s in reversed(range(len(self._checkpoints))):
            k0 = s * self.every
            k1 = min(k0 + self.every, self.nsteps)
            self._load(self._checkpoints[s])
            snaps = []
            for k in range(k0, k1):
                snaps.append(self._snapshot())
                self._step(k)
            for k in reversed(range(k0, k1)):
                self._step_adjoint(k, self.starts_init and k == 0, snaps[k - k0], bars)
        self.mesh.synchronize()
        return bars


class Barotropic_Adjoint(_Run_Adjoint):
    """Gradient of a function of the final state of an ``nsteps`` barotropic
    run with respect to the initial spectral vorticity (vor_p, vor_c).

    ``integrator`` supplies Δt, the Robert coefficient, the damping and
    whether the run starts with the initial (forward-Euler) step; it is not
    modified.  The grid fields of the initial state are recomputed from
    vor_c (grid_vor = T^-1 vor_c, (u, v) = UV_Grid_From_Vor_Div!(vor_c, 0)),
    so the run is a function of the spectral state alone.

        adj = Barotropic_Adjoint(mesh, atmo_data, integrator, checkpoint_every=8)
        vor_p, vor_c = adj.forward(vor_c0, vor_p0, nsteps)
        g_p, g_c = adj.backward(vor_p_bar, vor_c_bar)   # cotangents of the final state
    """

    STATE = ("spe_vor_p", "spe_vor_c")

    def __init__(self, mesh, atmo_data, integrator, checkpoint_every=8):
        super().__init__(mesh, atmo_data, integrator, checkpoint_every)

    def _load_grids(self):
        from .mesh import Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div
        d = self._d
        Trans_Spherical_To_Grid(self.mesh, d.spe_vor_c, d.grid_vor)
        UV_Grid_From_Vor_Div(self.mesh, d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)

    def _step(self, k):
        from .dynamics import Barotropic_Dynamics
        Barotropic_Dynamics(self.mesh, self.atmo, self._d, self._integrator(k))

    def _snapshot(self):
        d = self._d
        return d.grid_vor.clone(), d.grid_u_c.clone(), d.grid_v_c.clone()

    def _step_adjoint(self, k, init, snap, bars):
        zeta, u, v = snap
        barotropic_step_adjoint(self._integrator(k), init, zeta, u, v, self.atmo.omega, bars[0],
                                bars[1])

    def forward(self, vor_c, vor_p, nsteps):
        """Runs nsteps steps from (vor_p, vor_c) ([1, ns+1, N+1] device tensors),
        keeping a checkpoint every ``checkpoint_every`` steps; returns copies
        of the final (vor_p, vor_c)."""
        return self._run((vor_p, vor_c), nsteps)

    def backward(self, vor_p_bar, vor_c_bar):
        """Cotangents of the final (vor_p, vor_c) -> those of the initial state."""
        return self._reverse((vor_p_bar, vor_c_bar))


class Shallow_Water_Adjoint(_Run_Adjoint):
    """Gradient of a function of the final state of an ``nsteps``
    shallow-water run (Shallow_Water_Physics! then Shallow_Water_Dynamics!
    each step, exp/Shallow_Water) with respect to the initial spectral state
    (vor_p, vor_c, div_p, div_c, h_p, h_c).  The grid fields of the initial
    state are rebuilt from the spectral state (vor, div, h by the inverse
    transform, u, v by UV_Grid_From_Vor_Div!) and the height the advection
    term differentiates is h_c, as on every step after the first.

        adj = Shallow_Water_Adjoint(mesh, atmo, integrator, h_0, kappa_m, kappa_t, h_eq)
        final = adj.forward(state, nsteps)          # state = (vor_p, vor_c, div_p, div_c, h_p, h_c)
        grads = adj.backward(final_bars)            # same order
    """

    STATE = ("spe_vor_p", "spe_vor_c", "spe_div_p", "spe_div_c", "spe_lnps_p", "spe_lnps_c")

    def __init__(self, mesh, atmo_data, integrator, h_0, kappa_m, kappa_t, h_eq,
                 checkpoint_every=8):
        super().__init__(mesh, atmo_data, integrator, checkpoint_every)
        self.h_0, self.kappa_m, self.kappa_t = float(h_0), float(kappa_m), float(kappa_t)
        mesh._nlev(h_eq, "grid", True)
        self.h_eq = h_eq

    def _load_grids(self):
        from .mesh import Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div
        d, M = self._d, self.mesh
        d.spe_lnps_n.copy_(d.spe_lnps_c)
        Trans_Spherical_To_Grid(M, d.spe_vor_c, d.grid_vor)
        Trans_Spherical_To_Grid(M, d.spe_div_c, d.grid_div)
        Trans_Spherical_To_Grid(M, d.spe_lnps_c, d.grid_ps_c)
        UV_Grid_From_Vor_Div(M, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)

    def _step(self, k):
        from .dynamics import Shallow_Water_Physics, Shallow_Water_Dynamics
        Shallow_Water_Physics(self._d, self.kappa_m, self.kappa_t, self.h_eq)
        Shallow_Water_Dynamics(self.mesh, self.atmo, self.h_0, self._d, self._integrator(k))

    def _snapshot(self):
        d = self._d
        return (tuple(t.clone() for t in (d.grid_vor, d.grid_div, d.grid_ps_c, d.grid_u_c,
                                          d.grid_v_c)), d.spe_lnps_n.clone())

    def _step_adjoint(self, k, init, snap, bars):
        grids, spe_h = snap
        sw_step_adjoint(self._integrator(k), init, grids, spe_h, self.atmo.omega, self.kappa_m,
                        self.kappa_t, self.h_0, bars)

    def forward(self, state, nsteps):
        return self._run(tuple(state), nsteps)

    def backward(self, bars):
        return self._reverse(tuple(bars))
