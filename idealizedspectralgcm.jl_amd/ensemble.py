"""Batched ensembles of independent shallow-water instances.

No reference counterpart: the reference steps one model at a time.  All
instances of one truncation are the "levels" of one mesh (the shallow-water
step has no coupling between levels, so every kernel treats them as
independent columns of one batch); the meshes of the different truncations
share one HIP stream and one step of the whole mixed ensemble is captured
into one graph.  Per-instance results are identical to running each instance
alone, because every kernel computes each level-field with the same
arithmetic whatever the batch size.
"""
import ctypes

import numpy as np
import torch

from ._lib import call
from .dyn_data import Atmo_Data, Dyn_Data
from .dynamics import (Shallow_Water_Step, sw_closing_copy, sw_forward_batch, sw_grid_pass,
                       sw_inverse_batch, sw_spectral_pass)
from .graph import Step_Graph
from .mesh import (_ptr, Spectral_Spherical_Mesh, Trans_Spherical_To_Grid, UV_Grid_From_Vor_Div,
                   Vor_Div_From_Grid_UV)
from .plan import FORWARD, INVERSE, Transform_Plan
from .time_integrator import Filtered_Leapfrog, Update_Init_Step

RADIUS, OMEGA = 6371.0e3, 7.292e-5


def random_vorticity(N, count, rng, amplitude=1e-5):
    """Spectral vorticity with |ζ_mn| ~ N(0, σ_n), σ_n ∝ n^(-3/2) for 1 <= n <= N,
    m = 0 real, degree N+1 zero; normalised so the rms grid vorticity of an
    instance is about `amplitude` s^-1."""
    s = np.zeros((count, N + 2, N + 1), np.complex128)
    n = np.arange(N + 2, dtype=float)
    sig = np.zeros(N + 2)
    sig[1:N + 1] = n[1:N + 1] ** -1.5
    for m in range(N + 1):
        k = np.arange(m, N + 1)
        re = rng.standard_normal((count, k.size)) * sig[k]
        im = rng.standard_normal((count, k.size)) * sig[k] if m > 0 else 0.0
        s[:, k, m] = re + 1j * im
    # rms of the grid field = sqrt(sum |c|^2 weights) with unit-normalised harmonics
    power = np.sum(np.abs(s[:, :, 0]) ** 2, axis=1) + 2 * np.sum(np.abs(s[:, :, 1:]) ** 2, axis=(1, 2))
    s *= (amplitude / np.sqrt(power))[:, None, None]
    return s


def sw_step_workload(mesh, count, seed=0):
    """The shallow-water step's two transform batches at `count` instances
    (forward: u, v -> vor, div plus two plain arrays; inverse: seven arrays
    with the u cos θ, v cos θ and gradient operators), on arrays of their
    own, for the autotuner: (name, forward, inverse)."""
    from .mesh import (BATCH_DIVCOS, BATCH_VORDIV, INV_GRAD_LAMBDA, INV_GRAD_THETA, INV_UCOS,
                       INV_VCOS, Trans_Grid_To_Spherical_Batch, Trans_Spherical_To_Grid_Batch)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    grids = [torch.randn((count, mesh.nθ, mesh.nλ), dtype=torch.float64, device="cuda",
                         generator=gen) for _ in range(7)]
    specs = [mesh.spectral_zeros(count) for _ in range(4)]
    torch.cuda.synchronize()
    fwd_flags = [BATCH_VORDIV, 0, 0, 0]
    vor, div, h = specs[0], specs[1], specs[2]
    inv_in = [vor, div, h, vor, vor, h, h]
    inv_flags = [0, 0, 0, BATCH_DIVCOS | INV_UCOS, BATCH_DIVCOS | INV_VCOS,
                 BATCH_DIVCOS | INV_GRAD_LAMBDA, BATCH_DIVCOS | INV_GRAD_THETA]
    inv_b = [None, None, None, div, div, None, None]
    forward = lambda: Trans_Grid_To_Spherical_Batch(mesh, grids[:4], specs, fwd_flags)
    inverse = lambda: Trans_Spherical_To_Grid_Batch(mesh, inv_in, grids, inv_flags, snms_b=inv_b)
    return "sw-step", forward, inverse


class SW_Group:
    """`count` shallow-water instances at one truncation, stepped as one batch."""

    def __init__(self, N, nθ, count, dt, seed, precision="fp32", h_0=3.0e4, owner=None,
                 vor0=None, rotate=False):
        nλ = 2 * nθ
        self.rotate = bool(rotate)   # later steps: Time_Advance! by renaming p, c, n
        self.N, self.count, self.dt, self.h_0 = N, count, dt, h_0
        self.mesh = Spectral_Spherical_Mesh(N, N + 1, nθ, nλ, count, RADIUS, precision=precision)
        self._stream_owner = owner     # the library also keeps the owner alive
        if owner is not None:
            call("jgcm_mesh_share_stream", self.mesh.handle, owner.handle)
        m = self.mesh
        self.atmo = Atmo_Data("sw", nλ, nθ, count, False, False, False, False, m.sinθ, RADIUS, OMEGA)
        self.integ = Filtered_Leapfrog(0.04, 4, 1e-4, m.laplacian_eig, 0.5, dt, True, 0, 10 ** 9)
        d = self.d = Dyn_Data("sw", N, N + 1, nλ, nθ, count, mesh=m)
        if vor0 is None:
            vor0 = random_vorticity(N, count, np.random.default_rng(seed))
        d.spe_vor_c.copy_(torch.from_numpy(np.ascontiguousarray(vor0)))
        self.h_eq = torch.full((count, nθ, nλ), h_0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        # winds of the vorticity field, then linear balance ∇²Φ = ∇·(f ∇ψ)
        UV_Grid_From_Vor_Div(m, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
        zero = d.grid_d_full2
        zero.zero_()
        torch.cuda.synchronize()
        # (f v, -f u) = f ∇ψ for a nondivergent flow; its divergence is ∇·(f∇ψ)
        call("jgcm_vorticity_flux", m.handle, _ptr(zero), _ptr(d.grid_u_c), _ptr(d.grid_v_c),
             OMEGA, None, _ptr(d.grid_δu), _ptr(d.grid_δv), count, 0)
        Vor_Div_From_Grid_UV(m, d.grid_δu, d.grid_δv, d.spe_d1, d.spe_d2)
        m.synchronize()
        eig = torch.tensor(np.asarray(m.laplacian_eig).T, device="cuda")
        inv = torch.where(eig != 0, 1.0 / torch.where(eig != 0, eig, torch.ones_like(eig)),
                          torch.zeros_like(eig))
        phi = d.spe_d2 * inv[None]
        phi[:, 0, 0] += h_0                  # mean geopotential h_0 (P_0^0 = 1)
        torch.cuda.synchronize()
        # the SW step keeps h in the lnps/ps slots; one slot per instance
        d.spe_lnps_n = torch.zeros_like(d.spe_vor_c)
        d.spe_lnps_c = phi.contiguous()
        d.spe_lnps_p = phi.clone()
        d.spe_δlnps = torch.zeros_like(d.spe_vor_c)
        for f in ("grid_ps_n", "grid_ps_c", "grid_ps_p", "grid_δlnps"):
            setattr(d, f, torch.zeros_like(d.grid_u_c))
        d.spe_vor_p.copy_(d.spe_vor_c)
        torch.cuda.synchronize()
        Trans_Spherical_To_Grid(m, d.spe_lnps_c, d.grid_ps_c)
        Trans_Spherical_To_Grid(m, d.spe_vor_c, d.grid_vor)
        call("jgcm_copy", m.handle, _ptr(d.grid_ps_p), _ptr(d.grid_ps_c),
             d.grid_ps_c.numel() * 8)
        m.synchronize()
        self.steps = 0

    def step(self):
        """One shallow-water step of every instance of the group."""
        d, m = self.d, self.mesh
        Shallow_Water_Step(m, self.atmo, self.h_0, d, self.integ, 0.0, 0.0, self.h_eq,
                           carry_gradients=True, rotate=self.rotate)
        if self.steps == 0:
            Update_Init_Step(self.integ)
        self.steps += 1


class SW_Ensemble:
    """Mixed-truncation ensemble: groups = [(N, nθ, count, dt), ...]."""

    def __init__(self, groups, seed=0, precision="fp32", tuning=None, concurrent=False,
                 persistent=False, workgroups=-1, rotate=False):
        """`tuning`: None (built-in shapes) or an autotune.Tuning_Cache; each
        truncation's FFT and Legendre shapes (for its instance count) are then read from the
        cache, or tuned and stored when missing, before any step is captured.
        By default the truncations share the first group's stream and run one
        after another in the captured step; with `concurrent=True` each
        truncation keeps its own stream and the captured step forks them from
        the first group's stream and joins them back (jgcm_mesh_wait), so the
        truncations run as parallel branches of one graph.  The graph holds a
        library reference to every mesh it recorded, so no group's stream,
        events or scratch can be freed under it.

        `persistent=True` (one stream): the later steps run stage by stage over
        all truncations, and each transform direction is one plan
        (plan.Transform_Plan): the Legendre stage of every truncation in ONE
        launch from a cost-sorted work list dealt to `workgroups` persistent
        workgroups (< 0: as many as the device holds at once; 0: one per work
        item).  Each instance's results equal the other forms bit for bit.

        `rotate` (not with `persistent`, whose plans keep the arrays they were
        built for): the steps advance each group's state by renaming p, c, n
        (Shallow_Water_Step rotate) and a captured graph holds 3 steps."""
        from .autotune import Apply_Tuning
        assert not (concurrent and persistent), "the persistent form runs on one stream"
        assert not (persistent and rotate), "plans keep the arrays they were built for"
        self.rotate = bool(rotate)
        self.groups = []
        self.tuning = {}
        self.concurrent = bool(concurrent)
        self.persistent = bool(persistent)
        self.workgroups = int(workgroups)
        self._plan_fwd = self._plan_inv = None
        owner = None
        for i, (N, nθ, count, dt) in enumerate(groups):
            share = owner.mesh if (owner is not None and not self.concurrent) else None
            g = SW_Group(N, nθ, count, dt, seed + 1000 * i, precision=precision, owner=share,
                         rotate=rotate)
            if owner is None:
                owner = g
            if tuning is not None:
                e = Apply_Tuning(g.mesh, count, cache=tuning, tune_missing=True,
                                 workload=sw_step_workload(g.mesh, count))
                self.tuning[f"T{N}"] = None if e is None else \
                    (e["fold_ppw"], e["unfold_ppw"], e.get("fwd_tiles", 0), e.get("inv_tiles", 0))
            self.groups.append(g)
        self.owner = owner.mesh
        self.graph = None

    @property
    def instances(self):
        return sum(g.count for g in self.groups)

    def step(self):
        """One eager step of every group (each on its stream; see synchronize)."""
        if self.persistent and all(g.steps >= 1 for g in self.groups):
            self._stage_step()
            for g in self.groups:
                g.steps += 1
            return
        for g in self.groups:
            g.step()

    def plans(self):
        """The two transform plans of the persistent form (built on first use,
        after the first step, whose arrays they address)."""
        if self._plan_fwd is None:
            fwd = [(g.mesh, *sw_forward_batch(g.d)) for g in self.groups]
            inv = [(g.mesh, *sw_inverse_batch(g.d, True)) for g in self.groups]
            self._plan_fwd = Transform_Plan(FORWARD, fwd, self.workgroups)
            self._plan_inv = Transform_Plan(INVERSE, inv, self.workgroups)
        return self._plan_fwd, self._plan_inv

    def _stage_step(self):
        """One later step of every group, stage by stage (the stages of
        Shallow_Water_Step with carried gradients), the transforms as plans."""
        assert all(getattr(g.d, "_carried_sw", False) for g in self.groups), \
            "the persistent form follows a step that carried its gradients"
        fwd, inv = self.plans()
        for g in self.groups:
            sw_grid_pass(g.mesh, g.atmo, g.d, 0.0, 0.0, g.h_eq, True)
        fwd.execute()
        for g in self.groups:
            sw_spectral_pass(g.integ, g.h_0, g.d)
        inv.execute()
        for g in self.groups:
            sw_closing_copy(g.mesh, g.d)

    def synchronize(self):
        """Wait for every group's stream (graph launches complete on the owner's)."""
        for g in self.groups:
            g.mesh.synchronize()

    def capture(self):
        """Capture one later step of all groups into one graph (after the first step)."""
        assert all(g.steps >= 1 for g in self.groups), "take the first (init) step eagerly"
        for g in self.groups:
            g.mesh.reserve(7 * g.count)
        self.synchronize()
        others = [g for g in self.groups if g.mesh.stream() != self.owner.stream()]

        if self.persistent:
            self.plans()                           # created (and uploaded) before the capture
            self.graph = Step_Graph(self.owner, self._stage_step)
            return self.graph

        def body():
            for g in others:                       # fork: branches start after the owner
                call("jgcm_mesh_wait", g.mesh.handle, self.owner.handle)
            for g in self.groups:
                Shallow_Water_Step(g.mesh, g.atmo, g.h_0, g.d, g.integ, 0.0, 0.0, g.h_eq,
                                   carry_gradients=True, rotate=self.rotate)
            for g in others:                       # join: the graph ends after every branch
                call("jgcm_mesh_wait", self.owner.handle, g.mesh.handle)

        self.graph = Step_Graph(self.owner, body, repeat=3 if self.rotate else 1)
        return self.graph

    def launch(self):
        """One graph launch: 1 step of every group, 3 with rotated steps."""
        self.graph.launch()
        for g in self.groups:
            g.steps += self.graph.repeat
