"""Batched ensemble: every instance of a batched group evolves exactly like
the same instance run alone (bitwise), and a captured mixed-truncation step
replays like eager steps."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_concurrent_ensemble_is_switched_off(jgcm):
    from jgcm_amd.ensemble import SW_Ensemble
    with pytest.raises(RuntimeError, match="switched off"):
        SW_Ensemble([(21, 32, 1, 600)], concurrent=True)


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


@pytest.mark.parametrize("concurrent", [
    pytest.param(True, marks=pytest.mark.skip(
        reason="parallel-branch ensemble graph switched off: one replay segfaulted in "
               "hipGraphLaunch, cause not found")),
    False])
def test_mixed_ensemble_graph_matches_eager(jgcm, concurrent):
    """Replays of the captured step (truncations in sequence on one stream)
    equal eager steps bit for bit."""
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
