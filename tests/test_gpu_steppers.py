"""GPU parity of the barotropic and shallow-water steppers (host mirror over
the C ABI) against the FP64 oracle steppers, plus HIP-graph replay of a
step against eager execution.  Tolerance: 1e-11 relative after several
steps (FP64 path; only summation order differs)."""
import numpy as np
import pytest
import torch

from oracle import FilteredLeapfrog, SpectralSphericalMesh
from oracle import dynamics as odyn

pytestmark = pytest.mark.gpu

A, OMEGA = 6.371e6, 7.292e-5


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def to_dev(dev_d, ref_d, names):
    for n in names:
        getattr(dev_d, n).copy_(torch.from_numpy(np.ascontiguousarray(getattr(ref_d, n))))


STATE = ["spe_vor_c", "spe_vor_p", "spe_div_c", "spe_div_p", "spe_lnps_c", "spe_lnps_p",
         "grid_u_c", "grid_v_c", "grid_vor", "grid_div", "grid_ps_c", "grid_lnps"]


def rh4_setup(jgcm, N=42, nlat=64):
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    w = K = 7.848e-6
    mu = M.sin_theta[:, None]
    c = M.cos_theta[:, None]
    z = (2 * w * mu - K * mu * c ** 4 * 30 * np.cos(4 * M.lam_c[None, :]))[None]
    d = odyn.DynData("b", N, N + 1, 2 * nlat, nlat, 1)
    M.trans_grid_to_spherical(z, d.spe_vor_c)
    rng = np.random.default_rng(42)
    noise = 1e-8 * (rng.standard_normal(d.spe_vor_c.shape) + 1j * rng.standard_normal(d.spe_vor_c.shape))
    noise[:, :, 0] = noise[:, :, 0].real
    for m in range(N + 1):
        noise[:, :m, m] = 0
    noise[:, N + 1, :] = 0
    d.spe_vor_c += noise
    M.trans_spherical_to_grid(d.spe_vor_c, d.grid_vor)
    M.uv_grid_from_vor_div(d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)
    return M, dev, d


def test_rossby_haurwitz_builder_matches_oracle_state(jgcm):
    """Rossby_Haurwitz_Wave4 (device) equals the oracle's transforms of the same
    grid vorticity plus the same seeded spectral noise."""
    M, dev, d = rh4_setup(jgcm)
    dd = jgcm.Dyn_Data("b", 42, 43, 128, 64, 1, mesh=dev)
    jgcm.Rossby_Haurwitz_Wave4(dev, dd, noise=1e-8, seed=42)
    for f in ("spe_vor_c", "grid_vor", "grid_u_c", "grid_v_c", "spe_vor_p", "grid_u_p"):
        ref = getattr(d, f) if not f.endswith("_p") else getattr(d, f[:-2] + "_c")
        assert rel(getattr(dd, f).cpu().numpy(), ref) < 1e-12, f


def test_barotropic_steps_match_oracle(jgcm):
    M, dev, d = rh4_setup(jgcm)
    N, nlat = 42, 64
    atmo_r = odyn.AtmoData("b", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    Ir = FilteredLeapfrog(0.04, 4, 1.0 / (0.1 * 86400), M.laplacian_eig, 0.0, 1800, True, 0, 864000)
    dd = jgcm.Dyn_Data("b", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
    to_dev(dd, d, STATE)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A, OMEGA)
    Id = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), dev.laplacian_eig, 0.0, 1800, True,
                                0, 864000)
    np.testing.assert_allclose(Id.damping, Ir.damping, rtol=1e-14, atol=0)
    for i in range(6):
        odyn.barotropic_dynamics(M, atmo_r, d, Ir)
        jgcm.Barotropic_Dynamics(dev, atmo, dd, Id)
        if i == 0:
            Ir.update_init_step()
            jgcm.Update_Init_Step(Id)
    torch.cuda.synchronize()
    assert rel(dd.spe_vor_c.cpu().numpy(), d.spe_vor_c) < 1e-11
    assert rel(dd.spe_vor_p.cpu().numpy(), d.spe_vor_p) < 1e-11
    assert rel(dd.grid_u_c.cpu().numpy(), d.grid_u_c) < 1e-11
    assert rel(dd.grid_v_c.cpu().numpy(), d.grid_v_c) < 1e-11
    assert rel(dd.grid_vor.cpu().numpy(), d.grid_vor) < 1e-11
    with pytest.raises(AssertionError):
        jgcm.Update_Init_Step(Id)


def sw_setup(N, nlat):
    """exp/Shallow_Water/SW.jl:57-68 initial state (u = v = 0, grid_lnps = h_0,
    spe_lnps_c from it; grid_ps_c left at zero as the driver does) and the
    forcing h_eq of :70-77."""
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, A)
    d = odyn.DynData("sw", N, N + 1, 2 * nlat, nlat, 1)
    h_0 = 3.0e4
    d.grid_lnps[...] = h_0
    M.trans_grid_to_spherical(d.grid_lnps, d.spe_lnps_c)
    th, lam = M.theta_c, M.lam_c
    h_amp, h_lon, h_lat, h_width, h_itcz, itcz_w = 2.0e4, np.pi / 2, 25 * np.pi / 180, 15 * np.pi / 180, 1.0e5, 4 * np.pi / 180
    x2 = ((lam[None, :] - h_lon) / (2.0 * h_width)) ** 2
    y2 = ((th[:, None] - h_lat) / h_width) ** 2
    h_eq = (h_0 + h_amp * np.exp(-(x2 + y2)) + h_itcz * np.exp(-(th[:, None] / itcz_w) ** 2))[None]
    return M, d, h_0, np.ascontiguousarray(h_eq)


@pytest.mark.parametrize("N,nlat", [(42, 64), (31, 48)])
def test_shallow_water_steps_match_oracle(jgcm, N, nlat):
    M, d, h_0, h_eq = sw_setup(N, nlat)
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, 6.371e6)
    atmo_r = odyn.AtmoData("sw", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    Ir = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.5, 1200, True, 0, 691200)
    dd = jgcm.Dyn_Data("sw", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
    to_dev(dd, d, STATE)
    atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A, OMEGA)
    Id = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, dev.laplacian_eig, 0.5, 1200, True, 0, 691200)
    heq_d = torch.from_numpy(h_eq).cuda()
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
    for i in range(8):
        odyn.shallow_water_physics(d, km, kt, h_eq)
        odyn.shallow_water_dynamics(M, atmo_r, h_0, d, Ir)
        jgcm.Shallow_Water_Physics(dd, km, kt, heq_d)
        jgcm.Shallow_Water_Dynamics(dev, atmo, h_0, dd, Id)
        if i == 0:
            Ir.update_init_step()
            jgcm.Update_Init_Step(Id)
    torch.cuda.synchronize()
    for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_v_c", "grid_ps_c",
              "spe_δdiv", "spe_δlnps"):
        e = rel(getattr(dd, f).cpu().numpy(), getattr(d, f))
        assert e < 1e-11, (f, e)


def test_fused_shallow_water_step_equals_physics_and_dynamics(jgcm):
    """Shallow_Water_Step (the fused grid, height-tendency and spectral passes
    and the split time advance) equals Shallow_Water_Physics! +
    Shallow_Water_Dynamics! bit for bit in every array of Dyn_Data, over
    several steps (the first one forward) from the same state, with nonzero
    forcing."""
    N, nlat = 42, 64
    M, d, h_0, h_eq = sw_setup(N, nlat)
    runs = []
    for fused in (False, True):
        dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, 6.371e6)
        dd = jgcm.Dyn_Data("sw", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
        to_dev(dd, d, STATE)
        atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A,
                              OMEGA)
        Id = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, dev.laplacian_eig, 0.5, 1200, True, 0, 691200)
        heq_d = torch.from_numpy(h_eq).cuda()
        km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
        for i in range(4):
            if fused:
                jgcm.Shallow_Water_Step(dev, atmo, h_0, dd, Id, km, kt, heq_d)
            else:
                jgcm.Shallow_Water_Physics(dd, km, kt, heq_d)
                jgcm.Shallow_Water_Dynamics(dev, atmo, h_0, dd, Id)
            if i == 0:
                jgcm.Update_Init_Step(Id)
        torch.cuda.synchronize()
        runs.append({f: t.clone() for f, t in vars(dd).items() if isinstance(t, torch.Tensor)})
    assert len(runs[0]) > 30 and set(runs[0]) == set(runs[1])
    for f, v in runs[0].items():
        assert torch.equal(runs[1][f], v), f


def test_shallow_water_carried_gradients_equal_recomputed(jgcm):
    """Shallow_Water_Step carrying the gradient of the new h from one step's
    closing inverse batch to the next step's height tendency (eager steps,
    then a graph of two steps launched twice) leaves every array of Dyn_Data
    equal, bit for bit, to steps that transform h at their start."""
    N, nlat = 42, 64
    M, d, h_0, h_eq = sw_setup(N, nlat)
    runs = []
    for carry in (False, True):
        dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, 6.371e6)
        dd = jgcm.Dyn_Data("sw", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
        to_dev(dd, d, STATE)
        atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A,
                              OMEGA)
        Id = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, dev.laplacian_eig, 0.5, 1200, True, 0, 691200)
        heq_d = torch.from_numpy(h_eq).cuda()
        km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
        step = lambda: jgcm.Shallow_Water_Step(dev, atmo, h_0, dd, Id, km, kt, heq_d,
                                               carry_gradients=carry)
        step()
        jgcm.Update_Init_Step(Id)
        step()
        g = jgcm.Step_Graph(dev, step, nfields=7, repeat=2)
        g.launch()
        g.launch()
        g.close()
        step()
        torch.cuda.synchronize()
        runs.append({f: t.clone() for f, t in vars(dd).items() if isinstance(t, torch.Tensor)})
    assert len(runs[0]) > 30 and set(runs[0]) == set(runs[1])
    for f, v in runs[0].items():
        assert torch.equal(runs[1][f], v), f


def test_graph_replay_matches_eager(jgcm):
    """A captured SW step replayed 3 times equals 3 eager steps bit for bit;
    eager work and synchronisation happen between the replays."""
    N, nlat = 21, 32
    M, d, h_0, h_eq = sw_setup(N, nlat)
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, 6.371e6)
    atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A, OMEGA)
    heq_d = torch.from_numpy(h_eq).cuda()
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)

    def make():
        dd = jgcm.Dyn_Data("sw", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
        to_dev(dd, d, STATE)
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, dev.laplacian_eig, 0.5, 1200, True, 0, 691200)
        jgcm.Shallow_Water_Physics(dd, km, kt, heq_d)
        jgcm.Shallow_Water_Dynamics(dev, atmo, h_0, dd, I)
        jgcm.Update_Init_Step(I)
        return dd, I

    da, Ia = make()
    db, Ib = make()

    def step(dd, I):
        jgcm.Shallow_Water_Physics(dd, km, kt, heq_d)
        jgcm.Shallow_Water_Dynamics(dev, atmo, h_0, dd, I)

    g = jgcm.Step_Graph(dev, lambda: step(db, Ib), nfields=4)
    # capture does not execute: state b is still one step behind
    for k in range(3):
        step(da, Ia)
        g.launch()
        dev.synchronize()
        scratch = torch.ones(1000, device="cuda") * k   # eager work between replays
        torch.cuda.synchronize()
        for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_ps_c"):
            a = getattr(da, f).cpu().numpy()
            b = getattr(db, f).cpu().numpy()
            assert np.array_equal(a, b), (k, f)
    g.close()


def test_barotropic_rh4_ten_days_matches_oracle(jgcm):
    """The T42 barotropic case over its full 10 days (Δt = 1800 s, Robert 0.04,
    ∇⁸ damping with 0.1-day e-folding at n = 42): 480 device steps (the later
    ones replayed from one captured graph) against 480 oracle steps from the
    same wave-4 Rossby-Haurwitz state with seeded noise.  FP64; only the
    summation order differs, so the states agree to 1e-9 relative after 10 days
    (measured 1e-14)."""
    M, dev, d = rh4_setup(jgcm)
    N, nlat = 42, 64
    atmo_r = odyn.AtmoData("b", 2 * nlat, nlat, 1, False, False, False, False, M.sin_theta, A, OMEGA)
    Ir = FilteredLeapfrog(0.04, 4, 1.0 / (0.1 * 86400), M.laplacian_eig, 0.0, 1800, True, 0, 864000)
    dd = jgcm.Dyn_Data("b", N, N + 1, 2 * nlat, nlat, 1, mesh=dev)
    to_dev(dd, d, STATE)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, dev.sinθ, A, OMEGA)
    Id = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), dev.laplacian_eig, 0.0, 1800, True,
                                0, 864000)
    steps = 864000 // 1800
    for i in range(steps):
        odyn.barotropic_dynamics(M, atmo_r, d, Ir)
        if i == 0:
            Ir.update_init_step()
    jgcm.Barotropic_Dynamics(dev, atmo, dd, Id)
    jgcm.Update_Init_Step(Id)
    g = jgcm.Step_Graph(dev, lambda: jgcm.Barotropic_Dynamics(dev, atmo, dd, Id), nfields=4)
    for _ in range(steps - 1):
        g.launch()
    dev.synchronize()
    g.close()
    errs = {f: rel(getattr(dd, f).cpu().numpy(), getattr(d, f))
            for f in ("spe_vor_c", "spe_vor_p", "grid_vor", "grid_u_c", "grid_v_c")}
    print(errs)
    assert max(errs.values()) < 1e-9, errs
