"""Driver loops of the shallow-water experiments (exp/Shallow_Water/SW.jl:78-96
order: physics, dynamics, Update_Init_Step! after the first step), with the
later steps replayed from one captured HIP graph.

``Galewsky_Experiment`` is the unforced T85 shallow-water run of the Galewsky
jet (Δt = 600 s; Filtered_Leapfrog with Robert coefficient 0.04, ∇⁸ damping
1e-4 s^-1 at the truncation and implicit coefficient 0.5, the values of
exp/Shallow_Water/SW.jl:39-44).  ``run_days`` advances whole days and hands a
snapshot of the grid state to a callback at the end of each day.

``held_suarez_climate`` runs the Held-Suarez case (exp/HSt42/HS.jl order:
Spectral_Initialize_Fields!, then Atmosphere_Update! every step) and averages
its zonal-mean wind and temperature; ``jet_maxima`` locates the jets.
"""
import torch

from .dyn_data import Atmo_Data, Dyn_Data
from .dynamics import Shallow_Water_Step
from .graph import Step_Graph
from .initial_conditions import Galewsky_Jet
from .mesh import Spectral_Spherical_Mesh
from .time_integrator import Filtered_Leapfrog, Update_Init_Step

RADIUS, OMEGA = 6371.0e3, 7.292e-5


class Galewsky_Experiment:
    def __init__(self, N=85, nθ=128, dt=600, precision="fp64", seed=7, perturb=True,
                 robert_coef=0.04, damping_order=4, damping_coef=1.0e-4, implicit_coef=0.5,
                 end_time=6 * 86400, rotate=None):
        """`rotate` (default: when a day is a whole number of 3-step graphs):
        the later steps advance the state by renaming p, c, n
        (Shallow_Water_Step rotate) and replay as graphs of 3 steps."""
        nλ = 2 * nθ
        self.dt = int(dt)
        self.rotate = (86400 // self.dt) % 3 == 0 if rotate is None else bool(rotate)
        self.k = 3 if self.rotate else 1   # steps per graph
        m = self.mesh = Spectral_Spherical_Mesh(N, N + 1, nθ, nλ, 1, RADIUS, precision=precision)
        self.atmo = Atmo_Data("sw", nλ, nθ, 1, False, False, False, False, m.sinθ, RADIUS, OMEGA)
        self.integ = Filtered_Leapfrog(robert_coef, damping_order, damping_coef, m.laplacian_eig,
                                       implicit_coef, self.dt, True, 0, end_time)
        self.d = Dyn_Data("sw", N, N + 1, nλ, nθ, 1, mesh=m)
        self.info = Galewsky_Jet(m, self.d, seed=seed, perturb=perturb)
        self.h_0 = self.info["h_0"]
        # unforced: the physics call only clears the grid tendencies (κ = 0)
        self.h_eq = torch.zeros(1, nθ, nλ, dtype=torch.float64, device=m.device)
        self.steps = 0
        self.graph = None

    def _step(self):
        Shallow_Water_Step(self.mesh, self.atmo, self.h_0, self.d, self.integ, 0.0, 0.0, self.h_eq,
                           carry_gradients=True, rotate=self.rotate)

    def _prepare(self):
        """The first (forward) step eagerly, and with rotated steps two more,
        so that a whole number of 3-step graphs fills the first day; then the
        later steps are captured once."""
        if self.steps == 0:
            self._step()
            Update_Init_Step(self.integ)
            self.steps = 1
        while self.steps < self.k:
            self._step()
            self.steps += 1
        if self.graph is None:
            self.graph = Step_Graph(self.mesh, self._step, nfields=7, repeat=self.k)

    def snapshot(self):
        """(grid vorticity, divergence, geopotential, u) copies, after the stream drains."""
        self.mesh.synchronize()
        d = self.d
        out = tuple(t.clone() for t in (d.grid_vor, d.grid_div, d.grid_ps_c, d.grid_u_c))
        torch.cuda.synchronize()
        return out

    def run_days(self, days, on_day=None):
        """Advance `days` days (day k ends after k·86400/Δt steps); returns
        (device ms, steps timed): HIP events on the mesh stream around the
        graph replays; the eager first step, the capture and the snapshots
        are outside the timed region."""
        per_day = 86400 // self.dt
        stream = torch.cuda.ExternalStream(self.mesh.stream())
        ms, timed = 0.0, 0
        start_day = self.steps // per_day
        self._prepare()
        for day in range(start_day, start_day + days):
            target = (day + 1) * per_day
            self.mesh.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            n = target - self.steps
            assert n % self.k == 0, "a day is a whole number of graphs"
            for _ in range(n // self.k):
                self.graph.launch()
            e1.record(stream)
            self.mesh.synchronize()
            self.steps = target
            ms += e0.elapsed_time(e1)
            timed += n
            if on_day is not None:
                on_day(day + 1, self.snapshot())
        return ms, timed

    def close(self):
        if self.graph is not None:
            self.graph.close()
            self.graph = None


def relative_error(x, ref):
    """‖x - ref‖₂ / ‖ref‖₂ (the absolute norm where ref is exactly zero)."""
    n = float(torch.linalg.norm(ref))
    e = float(torch.linalg.norm(x - ref))
    return e / n if n > 0 else e


def precision_track(days=6, precisions=("split32", "bf16x3", "fp32"), seed=7, **kw):
    """Run the FP64 experiment and each reduced-precision one for `days` days;
    per day and field (vorticity, divergence, geopotential), the relative
    error against the FP64 run.  Returns {precision: {"ms_per_step", "errors":
    [[vor, div, h] per day 0..days]}} plus the FP64 entry."""
    out = {}
    snaps = {}
    for prec in ("fp64",) + tuple(precisions):
        e = Galewsky_Experiment(precision=prec, seed=seed, **kw)
        s = [e.snapshot()]
        ms, steps = e.run_days(days, lambda day, snap: s.append(snap))
        e.close()
        snaps[prec] = s
        out[prec] = {"ms_per_step": ms / steps, "steps": steps,
                     "sim_days_per_day": e.dt / (ms / steps * 1e-3)}
    ref = snaps["fp64"]
    for prec in precisions:
        out[prec]["errors"] = [[relative_error(a, b) for a, b in zip(sa[:3], sb[:3])]
                               for sa, sb in zip(snaps[prec], ref)]
    out["fp64"]["max_vor_by_day"] = [float(s[0].abs().max()) for s in ref]
    return out



HS_FORCING = {"σ_b": 0.7, "k_f": 1.0, "k_a": 1.0 / 40.0, "k_s": 1.0 / 4.0, "ΔT_y": 60.0,
              "Δθ_z": 10.0}


def held_suarez_climate(N=42, nθ=64, nd=20, dt=600, days=200, average_from=100,
                        precision="fp64", seed=5):
    """Held-Suarez run (exp/HSt42 driver: Spectral_Initialize_Fields! with seeded
    N(0, 0.1 K) temperature noise, then Atmosphere_Update! every Δt, later steps
    replayed from one captured graph) and its climate: zonal-mean u and T
    averaged once a day from `average_from` to `days`.  Returns the averages
    (numpy [nd][nθ]), the latitudes in degrees, σ at full levels, the device
    time per step and whether the state stayed finite."""
    import numpy as np
    from .graph import Step_Graph
    from .primitive import (Atmosphere_Update, Semi_Implicit_Solver, Spectral_Initialize_Fields,
                            Vert_Coordinate)
    nλ = 2 * nθ
    mesh = Spectral_Spherical_Mesh(N, N + 1, nθ, nλ, nd, RADIUS, precision=precision)
    vc = Vert_Coordinate(nλ, nθ, nd, "even_sigma", "simmons_and_burridge", "second_centered_wts",
                         1e5, mesh=mesh)
    atmo = Atmo_Data("hs", nλ, nθ, nd, True, True, False, False, mesh.sinθ, RADIUS, OMEGA)
    integ = Filtered_Leapfrog(0.04, 4, 1.15741e-4, mesh.laplacian_eig, 0.5, dt, True, 0,
                              days * 86400)
    si = Semi_Implicit_Solver(vc, atmo, integ, 1e5, np.full(nd, 300.0), mesh.wave_numbers)
    d = Dyn_Data("hs", N, N + 1, nλ, nθ, nd, mesh=mesh)
    d.grid_geopots.zero_()
    noise = torch.from_numpy(0.1 * np.random.default_rng(seed).standard_normal((nd, nθ, nλ)))
    Spectral_Initialize_Fields(mesh, atmo, vc, 1.0e5, 300.0, d.grid_geopots, d,
                               grid_t_noise=noise.to(mesh.device))
    # every step carries the next step's gradient transforms (the driver reads
    # the state between days but never changes it); when a day is a whole
    # number of 3-step graphs the state advances by renaming p, c, n (rotate)
    per_day = 86400 // dt
    rotate = per_day % 3 == 0
    step = lambda: Atmosphere_Update(mesh, atmo, vc, si, d, HS_FORCING, carry_gradients=True,
                                     rotate=rotate)
    step()
    Update_Init_Step(si)
    k = 3 if rotate else 1
    for _ in range(k - 1):   # the first day's eager steps: a whole number of graphs remains
        step()
    g = Step_Graph(mesh, step, nfields=8 * nd + 4, repeat=k)
    um = torch.zeros(nd, nθ, dtype=torch.float64, device=mesh.device)
    tm = torch.zeros_like(um)
    stream = torch.cuda.ExternalStream(mesh.stream())
    ms, timed, count = 0.0, 0, 0
    for day in range(days):
        n = per_day - k if day == 0 else per_day
        mesh.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n // k):
            g.launch()
        e1.record(stream)
        mesh.synchronize()
        ms += e0.elapsed_time(e1)
        timed += n
        if day >= average_from:
            um += d.grid_u_c.mean(dim=2)
            tm += d.grid_t_c.mean(dim=2)
            count += 1
            torch.cuda.synchronize()
    g.close()
    finite = bool(torch.isfinite(d.grid_t_c).all()) and bool(torch.isfinite(d.grid_u_c).all())
    return {"u_zonal": (um / max(count, 1)).cpu().numpy(),
            "t_zonal": (tm / max(count, 1)).cpu().numpy(),
            "lat_deg": np.degrees(np.asarray(mesh.θc)),
            "sigma": (np.arange(nd) + 0.5) / nd,
            "ms_per_step": ms / timed, "finite": finite}


def jet_maxima(climate):
    """(speed, σ, latitude) of the largest zonal-mean u in each hemisphere."""
    import numpy as np
    u, lat, sig = climate["u_zonal"], climate["lat_deg"], climate["sigma"]
    out = {}
    for name, sel in (("south", lat < 0), ("north", lat > 0)):
        sub = u[:, sel]
        k, j = np.unravel_index(np.argmax(sub), sub.shape)
        out[name] = (float(sub[k, j]), float(sig[k]), float(lat[sel][j]))
    return out


def _hs_model(N, nθ, nd, dt, precision, days):
    import numpy as np
    from .primitive import Semi_Implicit_Solver, Vert_Coordinate
    nλ = 2 * nθ
    mesh = Spectral_Spherical_Mesh(N, N + 1, nθ, nλ, nd, RADIUS, precision=precision)
    vc = Vert_Coordinate(nλ, nθ, nd, "even_sigma", "simmons_and_burridge", "second_centered_wts",
                         1e5, mesh=mesh)
    atmo = Atmo_Data("hs", nλ, nθ, nd, True, True, False, False, mesh.sinθ, RADIUS, OMEGA)
    integ = Filtered_Leapfrog(0.04, 4, 1.15741e-4, mesh.laplacian_eig, 0.5, dt, True, 0,
                              days * 86400)
    si = Semi_Implicit_Solver(vc, atmo, integ, 1e5, np.full(nd, 300.0), mesh.wave_numbers)
    d = Dyn_Data("hs", N, N + 1, nλ, nθ, nd, mesh=mesh)
    return mesh, vc, atmo, si, d


def _energy_enstrophy(mesh, d):
    """Mass-weighted global mean kinetic energy (u² + v²)/2 and area-weighted
    global mean enstrophy ζ²/2 over all levels (even σ: equal Δb)."""
    import numpy as np
    w = torch.tensor(np.asarray(mesh.wts), dtype=torch.float64, device=mesh.device)[None, :, None]
    ps = d.grid_ps_c
    ke = (0.5 * (d.grid_u_c ** 2 + d.grid_v_c ** 2) * ps * w).sum() / (ps * w).sum() / d.grid_u_c.shape[0]
    en = (0.5 * d.grid_vor ** 2 * w).sum() / (w.sum() * d.grid_vor.shape[0] * d.grid_vor.shape[2])
    return float(ke), float(en)


def precision_report(N=85, nθ=128, nd=20, dt=300, spinup_days=20, days=10,
                     precisions=("split32", "bf16x3", "fp32"), seed=5):
    """Errors of the reduced-precision Legendre stages on the Held-Suarez case,
    from one FP64 state spun up for `spinup_days`: per field, the relative L2
    error of one step's spectral tendencies (δζ, δD, δT, δ ln ps after the
    semi-implicit correction) against the FP64 step; after `days` more days, the
    relative difference of the global kinetic energy and enstrophy from the
    FP64 run and each run's relative change over those days.  The reduced
    precisions accumulate the Legendre sums in FP32, so every spectral
    coefficient carries an error of ~1e-7 of its field's norm; the small
    high-degree coefficients carry large relative errors, which ∇² (weight
    n(n+1)) and the semi-implicit solve bring out in δD (reported also against
    the norm of ∇²(Φ + KE), its largest explicit term)."""
    import numpy as np
    from .graph import Step_Graph
    from .primitive import Atmosphere_Update, Spectral_Initialize_Fields
    total = spinup_days + days + 1
    mesh, vc, atmo, si, d = _hs_model(N, nθ, nd, dt, "fp64", total)
    d.grid_geopots.zero_()
    noise = torch.from_numpy(0.1 * np.random.default_rng(seed).standard_normal((nd, nθ, 2 * nθ)))
    Spectral_Initialize_Fields(mesh, atmo, vc, 1.0e5, 300.0, d.grid_geopots, d,
                               grid_t_noise=noise.to(mesh.device))
    Atmosphere_Update(mesh, atmo, vc, si, d, HS_FORCING)
    Update_Init_Step(si)
    g = Step_Graph(mesh, lambda: Atmosphere_Update(mesh, atmo, vc, si, d, HS_FORCING),
                   nfields=5 * nd + 1)
    for _ in range(spinup_days * 86400 // dt - 1):
        g.launch()
    mesh.synchronize()
    g.close()
    names = [k for k, v in vars(d).items() if isinstance(v, torch.Tensor)]
    state = {k: getattr(d, k).clone() for k in names}
    torch.cuda.synchronize()
    tend = ("spe_δvor", "spe_δdiv", "spe_δt", "spe_δlnps", "spe_energy")
    runs = {}
    for prec in ("fp64",) + tuple(precisions):
        m2, vc2, atmo2, si2, d2 = _hs_model(N, nθ, nd, dt, prec, total)
        for k in names:
            getattr(d2, k).copy_(state[k])
        Update_Init_Step(si2)                    # later steps: 2Δt leapfrog and wave matrices
        torch.cuda.synchronize()
        e0 = _energy_enstrophy(m2, d2)
        Atmosphere_Update(m2, atmo2, vc2, si2, d2, HS_FORCING)
        m2.synchronize()
        t1 = {k: getattr(d2, k).clone() for k in tend}
        g2 = Step_Graph(m2, lambda: Atmosphere_Update(m2, atmo2, vc2, si2, d2, HS_FORCING),
                        nfields=5 * nd + 1)
        for _ in range(days * 86400 // dt - 1):
            g2.launch()
        m2.synchronize()
        g2.close()
        torch.cuda.synchronize()
        runs[prec] = (t1, e0, _energy_enstrophy(m2, d2))
    ref_t, ref_e0, ref_e1 = runs["fp64"]
    out = {"config": f"Held-Suarez T{N} L{nd}, dt {dt} s, from day {spinup_days} (FP64 spin-up)",
           "fp64": {"kinetic_energy": ref_e1[0], "enstrophy": ref_e1[1],
                    "kinetic_energy_change": ref_e1[0] / ref_e0[0] - 1.0,
                    "enstrophy_change": ref_e1[1] / ref_e0[1] - 1.0}}
    for prec in precisions:
        t1, e0, e1 = runs[prec]
        errs = {k.replace("spe_", ""): float(torch.linalg.norm(t1[k] - ref_t[k]) /
                                             torch.linalg.norm(ref_t[k])) for k in tend[:4]}
        # δD's error against its largest explicit term, ∇²(Φ + KE) (left in spe_energy)
        dd_vs_lap = float(torch.linalg.norm(t1["spe_δdiv"] - ref_t["spe_δdiv"]) /
                          torch.linalg.norm(ref_t["spe_energy"]))
        out[prec] = {"tendency_rel_err": errs, "δdiv_err_vs_laplacian_of_energy": dd_vs_lap,
                     "kinetic_energy_rel_diff": e1[0] / ref_e1[0] - 1.0,
                     "enstrophy_rel_diff": e1[1] / ref_e1[1] - 1.0,
                     "kinetic_energy_change": e1[0] / e0[0] - 1.0,
                     "enstrophy_change": e1[1] / e0[1] - 1.0}
    out["days_after_spinup"] = days
    return out
