"""Batched ensemble: every instance of a batched group evolves exactly like
the same instance run alone (bitwise), and a captured mixed-truncation step
replays like eager steps."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_batched_instances_equal_single_runs(jgcm):
    from jgcm_amd.ensemble import SW_Group, random_vorticity
    N, nlat, count, dt = 21, 32, 4, 600
    vor0 = random_vorticity(N, count, np.random.default_rng(7))
    g = SW_Group(N, nlat, count, dt, 0, precision="fp64", vor0=vor0)
    singles = [SW_Group(N, nlat, 1, dt, 0, precision="fp64", vor0=vor0[i:i + 1]) for i in range(count)]
    for _ in range(5):
        g.step()
        for s in singles:
            s.step()
    g.mesh.synchronize()
    for s in singles:
        s.mesh.synchronize()
    for i, s in enumerate(singles):
        for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_ps_c"):
            a = getattr(g.d, f)[i].cpu().numpy()
            b = getattr(s.d, f)[0].cpu().numpy()
            assert np.array_equal(a, b), (i, f)
    # the state is moving and finite
    assert np.isfinite(g.d.grid_u_c.cpu().numpy()).all()
    assert float(g.d.grid_u_c.abs().max()) > 0.1


@pytest.mark.parametrize("concurrent", [False, True])
def test_mixed_ensemble_graph_matches_eager(jgcm, concurrent):
    """Replays of the captured step (truncations in sequence on one stream, or
    as parallel branches on their own streams) equal eager steps bit for
    bit."""
    from jgcm_amd.ensemble import SW_Ensemble
    groups = [(21, 32, 3, 600), (42, 64, 2, 600), (85, 128, 2, 300)]
    A = SW_Ensemble(groups, seed=1, precision="fp32", concurrent=concurrent)
    B = SW_Ensemble(groups, seed=1, precision="fp32", concurrent=concurrent)
    A.step()
    B.step()
    B.capture()
    for k in range(3):
        A.step()
        B.launch()
        A.synchronize()
        B.owner.synchronize()
        for ga, gb in zip(A.groups, B.groups):
            for f in ("spe_vor_c", "spe_lnps_c", "grid_u_c"):
                assert np.array_equal(getattr(ga.d, f).cpu().numpy(), getattr(gb.d, f).cpu().numpy()), (k, f)


# The mixed ensemble of the benchmark: 64 x T21, 128 x T42, 64 x T85.
FULL = [(21, 32, 64, 1800), (42, 64, 128, 1200), (85, 128, 64, 600)]


def _oracle_instance(g, i):
    """The oracle's shallow-water model holding instance i of group g."""
    from oracle import FilteredLeapfrog, SpectralSphericalMesh
    from oracle import dynamics as odyn
    from jgcm_amd.ensemble import OMEGA, RADIUS
    N, nθ = g.N, g.mesh.nθ
    M = SpectralSphericalMesh(N, N + 1, nθ, 2 * nθ, 1, RADIUS)
    atmo = odyn.AtmoData("sw", 2 * nθ, nθ, 1, False, False, False, False, M.sin_theta, RADIUS, OMEGA)
    I = FilteredLeapfrog(0.04, 4, 1e-4, M.laplacian_eig, 0.5, g.dt, True, 0, 10 ** 9)
    d = odyn.DynData("sw", N, N + 1, 2 * nθ, nθ, 1)
    for name, val in vars(d).items():
        t = getattr(g.d, name, None)
        if isinstance(val, np.ndarray) and isinstance(t, torch.Tensor) and t.dim() == val.ndim \
                and t.shape[0] == g.count and tuple(t.shape[1:]) == val.shape[1:]:
            val[...] = t[i:i + 1].cpu().numpy()
    return M, atmo, I, d


def test_full_mixed_ensemble_sampled_instances_match_oracle(jgcm):
    """All 256 instances stepped together (FP64 Legendre): two instances of
    each truncation, the first and the last of its batch, follow the oracle's
    shallow-water model (Shallow_Water_Dynamics.jl:88-179) to 1e-11 over 5
    steps, the first step included."""
    from oracle import dynamics as odyn
    from jgcm_amd.ensemble import SW_Ensemble
    E = SW_Ensemble(FULL, seed=4, precision="fp64")
    assert E.instances == 256
    E.synchronize()
    picks = [(g, i, _oracle_instance(g, i)) for g in E.groups for i in (0, g.count - 1)]
    steps = 5
    for _ in range(steps):
        E.step()
    E.synchronize()
    for g, i, (M, atmo, I, d) in picks:
        h_eq = np.full((1, g.mesh.nθ, g.mesh.nλ), g.h_0)
        for k in range(steps):
            odyn.shallow_water_physics(d, 0.0, 0.0, h_eq)
            odyn.shallow_water_dynamics(M, atmo, g.h_0, d, I)
            if k == 0:
                I.update_init_step()
        for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_v_c", "grid_ps_c",
                  "grid_vor"):
            got = getattr(g.d, f)[i].cpu().numpy()
            ref = getattr(d, f)[0]
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < 1e-11, (f"T{g.N}", i, f, err)


def test_full_mixed_ensemble_graph_replays_equal_eager_steps(jgcm):
    """The benchmark's captured step of all 256 instances (FP32 Legendre,
    FP64 state): three replays equal three eager steps bit for bit and the
    state stays finite."""
    from jgcm_amd.ensemble import SW_Ensemble
    A = SW_Ensemble(FULL, seed=11, precision="fp32")
    B = SW_Ensemble(FULL, seed=11, precision="fp32")
    A.step()
    B.step()
    B.capture()
    for k in range(3):
        A.step()
        B.launch()
        A.synchronize()
        B.owner.synchronize()
        for ga, gb in zip(A.groups, B.groups):
            for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_ps_c"):
                a, b = getattr(ga.d, f), getattr(gb.d, f)
                assert torch.equal(a, b), (k, f"T{ga.N}", f)
                assert bool(torch.isfinite(b).all())
    B.graph.close()
