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
import ctypes

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


def trans_grid_to_spherical_adjoint_batch(mesh, snms, pfields):
    """trans_grid_to_spherical_adjoint of up to 12 arrays in two launches."""
    k = len(snms)
    assert len(pfields) == k and 1 <= k <= 8
    nlev = []
    for s, g in zip(snms, pfields):
        n = mesh._nlev(s, "spec", True)
        assert mesh._nlev(g, "grid", True) == n
        nlev.append(n)
    P = ctypes.c_void_p
    call("jgcm_trans_grid_to_spherical_adjoint_batch", mesh.handle, k,
         (P * k)(*[s.data_ptr() for s in snms]), (P * k)(*[g.data_ptr() for g in pfields]),
         (ctypes.c_int * k)(*nlev))


def trans_spherical_to_grid_adjoint_batch(mesh, pfields, snms, divcos=None, accumulate=False):
    """trans_spherical_to_grid_adjoint of up to 12 arrays in two launches;
    ``divcos[k]``: array k is divided by cos θ first; ``accumulate``: the
    results are added to ``snms`` (which must not overlap)."""
    k = len(pfields)
    assert len(snms) == k and 1 <= k <= 8
    nlev = []
    for g, s in zip(pfields, snms):
        n = mesh._nlev(g, "grid", True)
        assert mesh._nlev(s, "spec", True) == n
        nlev.append(n)
    P = ctypes.c_void_p
    flags = (ctypes.c_int * k)(*[1 if (divcos and divcos[i]) else 0 for i in range(k)])
    call("jgcm_trans_spherical_to_grid_adjoint_batch", mesh.handle, k,
         (P * k)(*[g.data_ptr() for g in pfields]), (P * k)(*[s.data_ptr() for s in snms]),
         (ctypes.c_int * k)(*nlev), flags, 1 if accumulate else 0)


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


def damped_leapfrog_adjoint_batch(integrator, init_step, fields):
    """damped_leapfrog_adjoint of up to 8 fields in one launch; ``fields`` =
    [(p_bar_next, c_bar_next, p_bar, c_bar, g_bar, damp), ...]."""
    mesh = integrator.mesh
    k = len(fields)
    assert 1 <= k <= 8
    nlev = []
    for f in fields:
        n = mesh._nlev(f[0], "spec", True)
        for t in f[1:5]:
            assert mesh._nlev(t, "spec", True) == n
        nlev.append(n)
    P = ctypes.c_void_p
    arr = lambda j: (P * k)(*[f[j].data_ptr() for f in fields])
    call("jgcm_damped_leapfrog_adjoint_batch", integrator.handle, 1 if init_step else 0, k,
         (ctypes.c_int * k)(*[1 if f[5] else 0 for f in fields]), arr(0), arr(1), arr(2), arr(3),
         arr(4), (ctypes.c_int * k)(*nlev))


def damped_leapfrog_adjoint(integrator, init_step, p_bar_next, c_bar_next, p_bar, c_bar, g_bar,
                            damp=True):
    """Transpose of Compute_Spectral_Damping! + Filtered_Leapfrog! + Time_Advance!
    (Time_Integrator.jl:58-101): (p_bar', c_bar') -> (p_bar, c_bar, g_bar);
    ``damp=False`` for a field the step does not damp."""
    mesh = integrator.mesh
    n = mesh._nlev(p_bar_next, "spec", True)
    for t in (c_bar_next, p_bar, c_bar, g_bar):
        assert mesh._nlev(t, "spec", True) == n
    call("jgcm_damped_leapfrog_adjoint", integrator.handle, 1 if init_step else 0, 1 if damp else 0,
         _ptr(p_bar_next), _ptr(c_bar_next), _ptr(p_bar), _ptr(c_bar), _ptr(g_bar), n)


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
    """Checkpointed reverse sweep over a run of a stepper whose state is a
    tuple of Dyn_Data fields.  Subclasses name the state fields (``STATE``,
    Dyn_Data attribute names), rebuild the derived grid fields from the state
    (``_load_grids``), run one step (``_step``), snapshot what the step adjoint
    needs before and after the step (``_snapshot``, ``_snapshot_after``) and
    run the reverse step (``_step_adjoint``)."""

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
        self._d = Dyn_Data("adjoint", mesh.num_fourier, mesh.num_spherical, mesh.nλ, mesh.nθ,
                           self._levels(), mesh=mesh)
        self._checkpoints = []
        self.nsteps = 0

    def _levels(self):
        return 1

    def _integrator(self, k):
        return self._first if k == 0 else self._later

    def _load(self, state):
        for name, t in zip(self.STATE, state):
            getattr(self._d, name).copy_(t)
        self._load_grids()

    def _state(self):
        return tuple(getattr(self._d, name).clone() for name in self.STATE)

    def _check_state(self, state):
        for t in state:
            self.mesh._nlev(t, "spec", True)
            assert t.shape[0] == 1, "the state has one level"

    def _snapshot_after(self):
        return None

    def _run(self, state, nsteps):
        assert nsteps >= 1
        assert len(state) == len(self.STATE)
        self._check_state(state)
        self._checkpoints = []
        self.nsteps = int(nsteps)
        self._load(state)
        for k in range(self.nsteps):
            if k % self.every == 0:
                self._checkpoints.append(self._state())
            self._run_step(k)
        return self._state()

    def _run_step(self, k):
        """Step k of the forward run."""
        self._step(k)

    # ---- one recorded trajectory at a time: Run keeps each apply's own ----
    def _trajectory(self):
        """What a reverse sweep needs from the last forward run."""
        return {"checkpoints": self._checkpoints, "nsteps": self.nsteps}

    def _restore(self, traj):
        self._checkpoints = traj["checkpoints"]
        self.nsteps = traj["nsteps"]

    def _forward_state(self, state, nsteps):
        """forward() on a STATE-ordered tuple."""
        return self._run(tuple(state), nsteps)

    def _backward_state(self, bars):
        """backward() on STATE-ordered cotangents."""
        return self._reverse(tuple(bars))

    def _working_bars(self, bars):
        """The cotangent buffers the reverse steps update in place."""
        return tuple(b.clone() for b in bars)

    def _reverse(self, bars):
        assert self._checkpoints, "forward() first"
        assert len(bars) == len(self.STATE)
        bars = self._working_bars(bars)
        for s in reversed(range(len(self._checkpoints))):
            k0 = s * self.every
            k1 = min(k0 + self.every, self.nsteps)
            self._load(self._checkpoints[s])
            snaps = [self._recompute(k) for k in range(k0, k1)]
            for k in reversed(range(k0, k1)):
                before, after = snaps[k - k0]
                snap = before if after is None else (before, after)
                self._step_adjoint(k, self.starts_init and k == 0, snap, bars)
        self.mesh.synchronize()
        return bars

    def _recompute(self, k):
        """Step k of a checkpoint segment again, with the snapshots its
        reverse step needs: (before, after)."""
        before = self._snapshot()
        self._step(k)
        return before, self._snapshot_after()


class Run(torch.autograd.Function):
    """A whole run as a differentiable function (the counterpart of a
    ChainRulesCore ``rrule`` of the time loop):

        final = Run.apply(adjoint, nsteps, *state)

    ``adjoint`` is a Barotropic_Adjoint, Shallow_Water_Adjoint or
    Held_Suarez_Adjoint; ``state`` its STATE tensors.  The backward pass is
    the adjoint's checkpointed reverse sweep; gradients of complex tensors use
    torch's convention (real and imaginary parts, <a, b> = Re(conj(a) b))."""

    @staticmethod
    def forward(ctx, adjoint, nsteps, *state):
        ctx.adjoint = adjoint
        ctx.shapes = [(t.shape, t.dtype) for t in state]
        out = adjoint._forward_state(tuple(t.detach().contiguous() for t in state), int(nsteps))
        # the adjoint object keeps only its latest run: this apply's checkpoints
        # travel with ctx, so several applies can be differentiated in any order
        ctx.trajectory = adjoint._trajectory()
        return out

    @staticmethod
    def backward(ctx, *bars):
        adj = ctx.adjoint
        full = tuple(b.contiguous() if b is not None else
                     torch.zeros(shp, dtype=dt, device=adj.mesh.device)
                     for b, (shp, dt) in zip(bars, ctx.shapes))
        adj._restore(ctx.trajectory)
        return (None, None) + tuple(adj._backward_state(full))


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


class Held_Suarez_Adjoint(_Run_Adjoint):
    """Gradient of a function of the final state of an ``nsteps`` Held-Suarez
    run (Atmosphere_Update! each step, exp/HSt42) with respect to its initial
    state: the spectral (p, c) pairs of vor, div, T and ln ps and the grid
    p-fields u_p, v_p, T_p, ps_p that the forcing and Compute_Corrections_Init
    read (the previous step's unfiltered grid state).  The grid c-fields are
    rebuilt from the spectral c-fields (inverse transforms, u and v by
    UV_Grid_From_Vor_Div!, ps = exp(ln ps)), as every step leaves them.

        adj = Held_Suarez_Adjoint(mesh, atmo, vert_coord, semi_implicit, hs_params)
        final = adj.forward(state, nsteps)    # state: tensors in Held_Suarez_Adjoint.STATE order
        grads = adj.backward(final_bars)      # same order

    Needs a pure sigma coordinate (even sigma, mcm or v197).  The reverse
    step composes C entry points: the leapfrog transposes, the implicit
    correction's transpose from its per-n matrix (built once per solver by
    applying the forward to unit vectors), the transposed transforms and
    ``jgcm_hs_column_adjoint`` / ``jgcm_hs_corrections_adjoint`` for the grid
    columns and the global corrections (whose scalar algebra runs on the
    device from the seven correction scalars the forward step leaves).
    """

    STATE = ("spe_vor_p", "spe_vor_c", "spe_div_p", "spe_div_c", "spe_t_p", "spe_t_c",
             "spe_lnps_p", "spe_lnps_c", "grid_u_p", "grid_v_p", "grid_t_p", "grid_ps_p")

    def __init__(self, mesh, atmo_data, vert_coord, semi_implicit, physics_params,
                 checkpoint_every=4, use_graphs=True):
        """``use_graphs``: after one eager run of each kind of reverse step
        (first / later), the step is captured into a HIP graph and replayed;
        each replay reads the step's snapshot from fixed staging buffers (one
        batched copy) and updates fixed cotangent buffers.  Results equal the
        eager steps bit for bit."""
        from .primitive import Get_Topography, Semi_Implicit_Solver
        self._nd = vert_coord.nd
        super().__init__(mesh, atmo_data, semi_implicit.integrator, checkpoint_every)
        mass, energy = bool(atmo_data.do_mass_correction), bool(atmo_data.do_energy_correction)
        assert mass == energy, "the adjoint supports both corrections or neither"
        self._corrections = mass
        self.vc = vert_coord
        self.params = dict(physics_params)
        prev = getattr(mesh, "_semi_implicit", None)

        def solver(I):
            return Semi_Implicit_Solver(vert_coord, atmo_data, I, semi_implicit.ps_ref,
                                        semi_implicit.t_ref, semi_implicit.wave_numbers)
        self._si_first, self._si_later = solver(self._first), solver(self._later)
        mesh._semi_implicit = prev
        Get_Topography(self._d.grid_geopots)
        self._matrices = {}
        self.use_graphs = bool(use_graphs)
        self._graphs = {}
        self._graph_runs = {}
        self._stage = None
        self._bars_fixed = None
        self._pool = {}
        self._slots = None
        self._snap_count = 0
        self._scal_bar = torch.zeros(4, dtype=torch.float64, device=mesh.device)
        self._param_b = torch.zeros((4, mesh.nθ, mesh.nλ), dtype=torch.float64, device=mesh.device)
        self.parameter_gradients = {}

    def _levels(self):
        return self._nd

    def _solver(self, k):
        return self._si_first if k == 0 else self._si_later

    # ---- forward -------------------------------------------------------------
    def _check_state(self, state):
        M, nd = self.mesh, self._nd
        for name, t in zip(self.STATE, state):
            kind = "spec" if name.startswith("spe_") else "grid"
            lev = 1 if name in ("spe_lnps_p", "spe_lnps_c", "grid_ps_p") else nd
            assert M._nlev(t, kind, True) == lev, f"{name}: {lev} levels expected"

    def _load_grids(self):
        from .mesh import Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div
        from .primitive import Compute_Pressures_And_Heights
        d, M = self._d, self.mesh
        Trans_Spherical_To_Grid(M, d.spe_vor_c, d.grid_vor, batch=True)
        Trans_Spherical_To_Grid(M, d.spe_div_c, d.grid_div, batch=True)
        UV_Grid_From_Vor_Div(M, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c, batch=True)
        Trans_Spherical_To_Grid(M, d.spe_t_c, d.grid_t_c, batch=True)
        Trans_Spherical_To_Grid(M, d.spe_lnps_c, d.grid_lnps, batch=True)
        call("jgcm_map", M.handle, 0, d.grid_lnps.numel(), _ptr(d.grid_lnps), 0.0, 0.0,
             _ptr(d.grid_ps_c))
        Compute_Pressures_And_Heights(self.atmo, self.vc, d.grid_ps_c, d.grid_geopots, d.grid_t_c,
                                      d.grid_p_half, d.grid_Δp, d.grid_lnp_half, d.grid_p_full,
                                      d.grid_lnp_full, d.grid_z_full, d.grid_z_half)

    def _step(self, k):
        from .primitive import Atmosphere_Update
        self._last = self._solver(k)
        Atmosphere_Update(self.mesh, self.atmo, self.vc, self._last, self._d, self.params)

    # Snapshots of the recomputed steps go to preallocated buffers, one set per
    # step of a checkpoint segment (a segment's snapshots are consumed by its
    # reverse steps, in stream order, before the next segment overwrites them),
    # each filled by one batched copy.
    _AFTER = ("grid_u_c", "grid_v_c", "grid_t_c", "grid_ps_c")

    def _slot(self):
        if self._slots is None or len(self._slots) != self.every:
            self._slots = [None] * self.every
        i = self._snap_count % self.every
        if self._slots[i] is None:
            d = self._d
            self._slots[i] = (tuple(torch.empty_like(getattr(d, n)) for n in self.STATE),
                              torch.zeros(7, dtype=torch.float64, device=self.mesh.device),
                              tuple(torch.empty_like(getattr(d, n)) for n in self._AFTER))
        return self._slots[i]

    def _copy_into(self, dsts, srcs):
        P, n = ctypes.c_void_p, len(srcs)
        call("jgcm_copy_batch", self.mesh.handle, n, (P * n)(*[t.data_ptr() for t in dsts]),
             (P * n)(*[t.data_ptr() for t in srcs]),
             (ctypes.c_size_t * n)(*[t.numel() * t.element_size() for t in srcs]))

    def _snapshot(self):
        before = self._slot()[0]
        self._copy_into(before, [getattr(self._d, n) for n in self.STATE])
        return before

    def _recompute(self, k):
        """With graphs, the recomputed step and its two snapshots are one
        captured graph per (kind of step, snapshot buffer), replayed after
        one eager run; the Python side (snapshot counter, last solver) is
        advanced here, since a replay runs no Python.  Step 0 (once per
        sweep) stays eager."""
        if not self.use_graphs or k == 0:
            return super()._recompute(k)
        from .graph import Step_Graph
        i = self._snap_count % self.every
        key = ("forward", i, tuple(sorted((str(a), float(b)) for a, b in self.params.items())))
        runs = self._graph_runs.get(key, 0)
        self._graph_runs[key] = runs + 1
        if runs == 0:
            return super()._recompute(k)   # eager: allocates the snapshot buffers
        g = self._graphs.get(key)
        if g is None:
            count = self._snap_count
            g = self._graphs[key] = Step_Graph(self.mesh, lambda: _Run_Adjoint._recompute(self, k))
            self._snap_count = count       # the capture ran the Python side once
        g.launch()
        self._last = self._solver(k)
        before, sc, after = self._slots[i]
        self._snap_count += 1
        return before, (sc, after)

    def _snapshot_after(self):
        _, sc, after = self._slot()
        call("jgcm_hs_scalars_device", self._last.handle, _ptr(sc), 7)
        self._copy_into(after, [getattr(self._d, n) for n in self._AFTER])
        self._snap_count += 1
        return sc, after

    def forward(self, state, nsteps):
        self._init_lnps = state[7].clone()
        return self._run(tuple(state), nsteps)

    def _trajectory(self):
        t = super()._trajectory()
        t["init_lnps"] = self._init_lnps
        return t

    def _restore(self, traj):
        super()._restore(traj)
        self._init_lnps = traj["init_lnps"]

    def _forward_state(self, state, nsteps):
        return self.forward(state, nsteps)

    def _backward_state(self, bars):
        return self.backward(bars)

    def backward(self, bars):
        """Cotangents of the final state -> those of the initial state (returned)
        and of the forcing parameters k_f, k_a, k_s, ΔT_y (in
        ``parameter_gradients``, keyed like the HS parameter dict)."""
        self._param_b.zero_()
        out = self._reverse(tuple(bars))
        sums = self._param_b.cpu().numpy().reshape(4, -1).sum(axis=1)
        self.parameter_gradients = dict(zip(("k_f", "k_a", "k_s", "ΔT_y"), map(float, sums)))
        return out

    def initial_state_gradient(self, grads):
        """Gradient with respect to the spectral coefficients (vor, div, T, ln ps) of a
        state whose p-fields equal its c-fields and whose grid p-fields are its grid
        c-fields (the state Spectral_Initialize_Fields! leaves): the c and p
        cotangents added, and the grid p-field cotangents taken back through
        UV_Grid_From_Vor_Div!, the inverse transform and ps = exp(ln ps)."""
        from .mesh import Trans_Spherical_To_Grid
        M, nd, h = self.mesh, self._nd, self.mesh.handle
        gvp, gvc, gdp, gdc, gtp, gtc, glp, glc, gup, gvpg, gtpg, gpsp = grads
        out = [gvc.clone(), gdc.clone(), gtc.clone(), glc.clone()]

        def add(dst, src):
            call("jgcm_axpby", h, dst.numel() * dst.element_size() // 8, 1.0, _ptr(dst), 1.0,
                 _ptr(src), _ptr(dst))
        for dst, src in zip(out, (gvp, gdp, gtp, glp)):
            add(dst, src)
        sv, sd = M.spectral_zeros(nd), M.spectral_zeros(nd)
        uv_grid_from_vor_div_adjoint(M, gup, gvpg, sv, sd)
        add(out[0], sv)
        add(out[1], sd)
        st = M.spectral_zeros(nd)
        trans_spherical_to_grid_adjoint(M, gtpg, st)
        add(out[2], st)
        ln_g, ps_g, lg, sl = M.grid_zeros(1), M.grid_zeros(1), M.grid_zeros(1), M.spectral_zeros(1)
        Trans_Spherical_To_Grid(M, self._init_lnps, ln_g, batch=True)
        call("jgcm_map", h, 0, ln_g.numel(), _ptr(ln_g), 0.0, 0.0, _ptr(ps_g))
        call("jgcm_xmy", h, lg.numel(), 1.0, _ptr(ps_g), _ptr(gpsp), _ptr(lg), 0)
        trans_spherical_to_grid_adjoint(M, lg, sl)
        add(out[3], sl)
        M.synchronize()
        return tuple(out)

    # ---- reverse ---------------------------------------------------------------
    def _implicit_matrix(self, si):
        """[ns+1][2nd+1][7nd+3] matrix of ∇²(energy) + Implicit_Correction! per n,
        column i = the forward applied to the i-th unit input at m = 0."""
        key = id(si)
        if key in self._matrices:
            return self._matrices[key]
        M, nd = self.mesh, self._nd
        names = [("ddiv", nd), ("dlnps", 1), ("dt", nd), ("energy", nd), ("div_c", nd),
                 ("div_p", nd), ("lnps_c", 1), ("lnps_p", 1), ("t_c", nd), ("t_p", nd)]
        buf = {n: M.spectral_zeros(L) for n, L in names}
        nin, nout = 7 * nd + 3, 2 * nd + 1
        A = torch.zeros((M.num_spherical + 1, nout, nin), dtype=torch.float64, device=M.device)
        i = 0
        for name, L in names:
            for lev in range(L):
                for t in buf.values():
                    t.zero_()
                buf[name][lev, :, 0] = 1.0
                call("jgcm_hs_energy_implicit_correction", si.handle, *map(_ptr, (
                    buf["energy"], buf["div_c"], buf["div_p"], buf["lnps_c"], buf["lnps_p"],
                    buf["t_c"], buf["t_p"], buf["ddiv"], buf["dlnps"], buf["dt"])))
                M.synchronize()
                A[:, 0:nd, i] = buf["ddiv"][:, :, 0].real.T
                A[:, nd, i] = buf["dlnps"][0, :, 0].real
                A[:, nd + 1:, i] = buf["dt"][:, :, 0].real.T
                i += 1
        A = A.contiguous()
        self._matrices[key] = A
        return A

    def _working_bars(self, bars):
        if not self.use_graphs:
            return super()._working_bars(bars)
        fixed = self._bars_fixed
        if fixed is None or any(a.shape != b.shape or a.dtype != b.dtype for a, b in zip(fixed, bars)):
            self._close_graphs()
            fixed = self._bars_fixed = tuple(torch.empty_like(b) for b in bars)
        torch.cuda.synchronize()
        for a, b in zip(fixed, bars):
            a.copy_(b)
        torch.cuda.synchronize()
        return fixed

    def _reverse(self, bars):
        out = super()._reverse(bars)
        # with graphs the sweep works in fixed buffers that the next sweep reuses
        return tuple(t.clone() for t in out) if self.use_graphs else out

    def _close_graphs(self):
        """Drop the reverse-step graphs (they read the staging and cotangent
        buffers being replaced); the forward-step graphs stay valid."""
        for key in [q for q in self._graphs if q[0] == "reverse"]:
            self._graphs.pop(key).close()
        for key in [q for q in self._graph_runs if q[0] == "reverse"]:
            del self._graph_runs[key]

    def _step_adjoint(self, k, init, snap, bars):
        if not self.use_graphs:
            return self._step_adjoint_eager(k, init, snap, bars)
        from .graph import Step_Graph
        M = self.mesh
        st, (sc, nxt) = snap
        flat = list(st) + [sc] + list(nxt)
        if self._stage is None or any(a.shape != b.shape for a, b in zip(self._stage, flat)):
            self._close_graphs()
            self._stage = [torch.empty_like(t) for t in flat]
        stage = self._stage
        ptrs = lambda ts: (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
        call("jgcm_copy_batch", M.handle, len(flat), ptrs(stage), ptrs(flat),
             (ctypes.c_size_t * len(flat))(*[t.numel() * t.element_size() for t in flat]))
        staged = (tuple(stage[:len(st)]), (stage[len(st)], tuple(stage[len(st) + 1:])))
        # one graph per kind of step (solver, integrator, forcing parameters)
        key = ("reverse", k == 0, bool(init),
               tuple(sorted((str(a), float(b)) for a, b in self.params.items())))
        runs = self._graph_runs.get(key, 0)
        if runs == 0:
            # eager: fills the temporaries' pool, the implicit matrix and the scratch
            self._step_adjoint_eager(k, init, staged, bars)
        else:
            g = self._graphs.get(key)
            if g is None:
                g = self._graphs[key] = Step_Graph(
                    M, lambda: self._step_adjoint_eager(k, init, staged, bars))
            g.launch()
        self._graph_runs[key] = runs + 1

    def _step_adjoint_eager(self, k, init, snap, bars):
        from .mesh import (Trans_Spherical_To_Grid_Batch, BATCH_DIVCOS, INV_GRAD_LAMBDA,
                           INV_GRAD_THETA, INV_UCOS, INV_VCOS)
        from .primitive import _hs
        from .time_integrator import Get_Δt
        M, nd, atmo = self.mesh, self._nd, self.atmo
        h = M.handle
        st, (sc, nxt) = snap
        vor_p, vor_c, div_p, div_c, t_p, t_c, lnps_p, lnps_c, u_pg, v_pg, t_pg, ps_pg = st
        Bvp, Bvc, Bdp, Bdc, Btp, Btc, Blp, Blc, Bup, Bvpg, Btpg, Bpsp = bars
        si, I = self._solver(k), self._integrator(k)
        # temporaries: one zero-initialised buffer per allocation site, kept
        # across steps (the n-th request of a step always comes from the same
        # call, whose producer overwrites it), so no per-step fill is launched
        site = [0]

        def take(kind, nlev):
            key = (site[0], kind, nlev)
            site[0] += 1
            t = self._pool.get(key)
            if t is None:
                t = (M.spectral_zeros if kind == "s" else M.grid_zeros)(nlev)
                self._pool[key] = t
            return t

        def S(nlev):
            return take("s", nlev)

        def G(nlev):
            return take("g", nlev)

        def add(*pairs):            # dst += src for each (dst, src) pair, one launch
            P, n = ctypes.c_void_p, len(pairs)
            call("jgcm_add_batch", h, n, (P * n)(*[d.data_ptr() for d, _ in pairs]),
                 (P * n)(*[s.data_ptr() for _, s in pairs]),
                 (ctypes.c_size_t * n)(*[d.numel() * d.element_size() // 8 for d, _ in pairs]))

        def t2g_T_add(grids, targets):   # targets[k] += transpose of the inverse (grids[k])
            trans_spherical_to_grid_adjoint_batch(M, grids, targets, accumulate=True)

        # 1. mass and energy corrections (Spectral_Dynamics.jl:30-68); the scalar
        #    cotangents stay on the device (no wait for the stream)
        u_n, v_n, t_nc, ps_nc = nxt
        scal_bar = self._scal_bar          # stays zero without the corrections
        if self._corrections:
            ub, vb, tb, lb = G(nd), G(nd), G(nd), G(1)
            call("jgcm_hs_corrections_adjoint", si.handle, *map(_ptr, (
                sc, Btc, Blc, u_n, v_n, t_nc, ps_nc, scal_bar, ub, vb, tb, lb)))
            t2g_T_add([tb, lb], [Btc, Blc])
            sv, sd = S(nd), S(nd)
            uv_grid_from_vor_div_adjoint(M, ub, vb, sv, sd)
            add((Bvc, sv), (Bdc, sd))
        # 2. damping + leapfrog + Time_Advance (vor, div, T damped; ln ps not)
        Gv, Gd, Gt, Gl = S(nd), S(nd), S(nd), S(1)
        damped_leapfrog_adjoint_batch(I, init, [(Bvp, Bvc, Bvp, Bvc, Gv, True),
                                                (Bdp, Bdc, Bdp, Bdc, Gd, True),
                                                (Btp, Btc, Btp, Btc, Gt, True),
                                                (Blp, Blc, Blp, Blc, Gl, False)])
        # 3. ∇²(energy), δdiv -= energy, Implicit_Correction!
        out = S(7 * nd + 3)
        call("jgcm_implicit_adjoint", h, nd, _ptr(self._implicit_matrix(si)), _ptr(Gd), _ptr(Gl),
             _ptr(Gt), _ptr(out))
        o_ddiv, o_dl, o_dt, o_en = out[0:nd], out[nd:nd + 1], out[nd + 1:2 * nd + 1], out[2 * nd + 1:3 * nd + 1]
        base = 3 * nd + 1
        add((Bdc, out[base:base + nd]), (Bdp, out[base + nd:base + 2 * nd]),
            (Blc, out[base + 2 * nd:base + 2 * nd + 1]),
            (Blp, out[base + 2 * nd + 1:base + 2 * nd + 2]),
            (Btc, out[base + 2 * nd + 2:base + 3 * nd + 2]),
            (Btp, out[base + 3 * nd + 2:base + 4 * nd + 2]))
        # 4. forward transforms of δln ps, δT, Φ+KE and (δu, δv) -> (δζ, δD)
        dl_g, dt_g, en_g = G(1), G(nd), G(nd)
        trans_grid_to_spherical_adjoint_batch(M, [o_dl, o_dt, o_en], [dl_g, dt_g, en_g])
        du_g, dv_g = G(nd), G(nd)
        vor_div_from_grid_uv_adjoint(M, Gv, o_ddiv.contiguous(), du_g, dv_g)
        # 5. grid columns at the step's state
        #    (the forward step's two inverse batches: same kernels, same values)
        vor_g, div_g, u_g, v_g, t_g, ln_g, ps_c = G(nd), G(nd), G(nd), G(nd), G(nd), G(1), G(1)
        D = BATCH_DIVCOS
        Trans_Spherical_To_Grid_Batch(
            M, [vor_c, div_c, vor_c, vor_c, t_c, lnps_c], [vor_g, div_g, u_g, v_g, t_g, ln_g],
            [0, 0, D | INV_UCOS, D | INV_VCOS, 0, 0],
            snms_b=[None, None, div_c, div_c, None, None])
        call("jgcm_map", h, 0, ln_g.numel(), _ptr(ln_g), 0.0, 0.0, _ptr(ps_c))
        gl, gt, tl, tt = G(1), G(1), G(nd), G(nd)
        Trans_Spherical_To_Grid_Batch(M, [lnps_c, lnps_c, t_c, t_c], [gl, gt, tl, tt],
                                      [D | INV_GRAD_LAMBDA, D | INV_GRAD_THETA,
                                       D | INV_GRAD_LAMBDA, D | INV_GRAD_THETA])
        outs = [G(nd), G(nd), G(nd), G(nd), G(nd), G(nd), G(nd), G(1), G(1), G(nd), G(nd), G(nd),
                G(1), G(1)]
        div_b, u_b, v_b, t_b, vor_b, tl_b, tt_b, gl_b, gt_b, up_b, vp_b, tp_b, psp_b, psc_b = outs
        arr = lambda ts: (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
        hp = _hs(self.params)
        call("jgcm_hs_column_adjoint", si.handle, ctypes.byref(hp), float(Get_Δt(I)), _ptr(scal_bar),
             arr([div_g, u_g, v_g, t_g, vor_g, tl, tt, gl, gt, u_pg, v_pg, t_pg, ps_pg, ps_c]),
             arr([du_g, dv_g, dt_g, en_g, dl_g]), arr(outs), _ptr(self._param_b))
        # 6. the previous cotangents of the grid p-fields belong to this step's grid c-fields
        add((u_b, Bup), (v_b, Bvpg), (t_b, Btpg), (psc_b, Bpsp))
        cp = ((Bup, up_b), (Bvpg, vp_b), (Btpg, tp_b), (Bpsp, psp_b))
        P4 = ctypes.c_void_p * 4
        call("jgcm_copy_batch", h, 4, P4(*[d.data_ptr() for d, _ in cp]),
             P4(*[s.data_ptr() for _, s in cp]),
             (ctypes.c_size_t * 4)(*[s.numel() * s.element_size() for _, s in cp]))
        # 7. grid c-fields -> spectral c-fields
        lg = G(1)
        call("jgcm_xmy", h, lg.numel(), 1.0, _ptr(ps_c), _ptr(psc_b), _ptr(lg), 0)
        t2g_T_add([vor_b, div_b, t_b, lg], [Bvc, Bdc, Btc, Blc])
        sv, sd = S(nd), S(nd)
        uv_grid_from_vor_div_adjoint(M, u_b, v_b, sv, sd)
        add((Bvc, sv), (Bdc, sd))
        # gradients of ln ps and T (the inverse batch divided them by cos θ);
        # adjacent outputs, zeroed by one fill
        s_all = S(2 + 2 * nd)
        s_ll, s_lt, s_tl, s_tt = s_all[0:1], s_all[1:2], s_all[2:2 + nd], s_all[2 + nd:]
        trans_spherical_to_grid_adjoint_batch(M, [gl_b, gt_b, tl_b, tt_b],
                                              [s_ll, s_lt, s_tl, s_tt], divcos=[1, 1, 1, 1])
        call("jgcm_gradient_cos_adjoint", h, _ptr(s_ll), _ptr(s_lt), _ptr(Blc), 1, 1)
        call("jgcm_gradient_cos_adjoint", h, _ptr(s_tl), _ptr(s_tt), _ptr(Btc), 1, nd)
