"""Reverse-mode derivatives: the transform pair and a barotropic run.

The transforms are linear, so their vector-Jacobian products are the adjoint
transforms of the C ABI (weighted-transpose Legendre + conjugate FFT).  They
are exposed as ``torch.autograd.Function`` subclasses, the Python counterpart
of a ChainRulesCore ``rrule`` on the Julia side; gradients are with respect to
the real grid values and, for spectral arrays, with respect to the real and
imaginary parts (torch's convention for complex tensors, <a, b> = Re(conj(a) b)).

``Barotropic_Adjoint`` differentiates a whole run of ``Barotropic_Dynamics!``
(Barotropic_Dynamics.jl:2-46, with ``Update_Init_Step!`` after the first step
as in exp/Barotropic) with respect to its initial spectral state: a forward
pass that keeps the state every ``checkpoint_every`` steps, then a reverse
sweep that recomputes each segment forward (keeping its grid fields) and runs
``jgcm_barotropic_step_adjoint`` backwards through it.  Device memory is
O(nsteps / checkpoint_every + checkpoint_every) states.
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


def barotropic_step_adjoint(integrator, init_step, grid_vor, grid_u, grid_v, omega, p_bar, c_bar):
    """One reverse step of Barotropic_Dynamics!, in place on (p_bar, c_bar)."""
    mesh = integrator.mesh
    n = mesh._nlev(p_bar, "spec", True)
    assert mesh._nlev(c_bar, "spec", True) == n
    for t in (grid_vor, grid_u, grid_v):
        assert mesh._nlev(t, "grid", True) == n
    call("jgcm_barotropic_step_adjoint", integrator.handle, 1 if init_step else 0, _ptr(grid_vor),
         _ptr(grid_u), _ptr(grid_v), float(omega), _ptr(p_bar), _ptr(c_bar), n)


class Barotropic_Adjoint:
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

    def __init__(self, mesh, atmo_data, integrator, checkpoint_every=8):
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
        self._d = Dyn_Data("barotropic_adjoint", mesh.num_fourier, mesh.num_spherical, mesh.nλ,
                           mesh.nθ, 1, mesh=mesh)
        self._checkpoints = []
        self.nsteps = 0

    def _load(self, vor_p, vor_c):
        from .mesh import Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div
        d = self._d
        d.spe_vor_p.copy_(vor_p)
        d.spe_vor_c.copy_(vor_c)
        Trans_Spherical_To_Grid(self.mesh, d.spe_vor_c, d.grid_vor)
        UV_Grid_From_Vor_Div(self.mesh, d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)

    def _step(self, k):
        from .dynamics import Barotropic_Dynamics
        I = self._first if k == 0 else self._later
        Barotropic_Dynamics(self.mesh, self.atmo, self._d, I)

    def forward(self, vor_c, vor_p, nsteps):
        """Runs nsteps steps from (vor_p, vor_c) ([1, ns+1, N+1] device tensors),
        keeping a checkpoint every ``checkpoint_every`` steps; returns copies
        of the final (vor_p, vor_c)."""
        assert nsteps >= 1
        self.mesh._nlev(vor_c, "spec", True)
        self.mesh._nlev(vor_p, "spec", True)
        assert vor_c.shape[0] == 1 and vor_p.shape[0] == 1, "the barotropic state has one level"
        self._checkpoints = []
        self.nsteps = int(nsteps)
        self._load(vor_p, vor_c)
        d = self._d
        for k in range(self.nsteps):
            if k % self.every == 0:
                self._checkpoints.append((d.spe_vor_p.clone(), d.spe_vor_c.clone()))
            self._step(k)
        return d.spe_vor_p.clone(), d.spe_vor_c.clone()

    def backward(self, vor_p_bar, vor_c_bar):
        """Cotangents of the final (vor_p, vor_c) -> those of the initial state."""
        assert self._checkpoints, "forward() first"
        p_bar, c_bar = vor_p_bar.clone(), vor_c_bar.clone()
        d = self._d
        for s in reversed(range(len(self._checkpoints))):
            k0 = s * self.every
            k1 = min(k0 + self.every, self.nsteps)
            self._load(*self._checkpoints[s])
            grids = []
            for k in range(k0, k1):
                grids.append((d.grid_vor.clone(), d.grid_u_c.clone(), d.grid_v_c.clone()))
                self._step(k)
            for k in reversed(range(k0, k1)):
                zeta, u, v = grids[k - k0]
                init = self.starts_init and k == 0
                I = self._first if k == 0 else self._later
                barotropic_step_adjoint(I, init, zeta, u, v, self.atmo.omega, p_bar, c_bar)
        self.mesh.synchronize()
        return p_bar, c_bar
