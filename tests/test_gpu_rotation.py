"""Rotated Time_Advance! (Atmosphere_Update(..., rotate=True)): renaming the
state's p, c, n arrays instead of copying them leaves p and c of every
prognostic array, the diagnostic grid arrays and the deferred pressures and
heights equal, bit for bit, to the copying step (eager steps and graphs of
three steps); graphs whose steps do not bring the names back are refused and
leave the names as they were; a graph launched after an eager rotated step is
refused."""
import pytest
import torch

from tests.test_gpu_primitive import HS, build

pytestmark = pytest.mark.gpu

# arrays a rotated step leaves stale (the n of the renamed state)
N_FIELDS = {f"{f}_n" for f in ("spe_vor", "spe_div", "spe_lnps", "spe_t", "grid_u", "grid_v",
                               "grid_ps", "grid_t")}


def snapshot(dd):
    return {f: getattr(dd, f).clone() for f in list(vars(dd)) + ["grid_p_full", "grid_z_full"]
            if isinstance(getattr(dd, f, None), torch.Tensor) and f not in N_FIELDS
            and not f.startswith("_stored_")}


@pytest.mark.parametrize("N,nlat", [(42, 64), (85, 128)])
def test_rotated_steps_equal_copying_steps(jgcm, N, nlat):
    runs = []
    for rotate in (False, True):
        _, (dev, dvc, datmo, dsi, dd) = build(jgcm, N, nlat, 20, 600 if N == 42 else 300)
        step = lambda: jgcm.Atmosphere_Update(dev, datmo, dvc, dsi, dd, HS, carry_gradients=True,
                                              rotate=rotate)
        step()
        jgcm.Update_Init_Step(dsi)
        step()
        step()
        g = jgcm.Step_Graph(dev, step, nfields=8 * 20 + 4, repeat=3)
        g.launch()
        g.launch()
        g.close()
        step()
        dev.synchronize()
        runs.append(snapshot(dd))
    assert len(runs[0]) > 50 and set(runs[0]) == set(runs[1])
    for f, v in runs[0].items():
        assert torch.equal(runs[1][f], v), f


def test_rotated_graph_needs_whole_rotations(jgcm):
    _, (dev, dvc, datmo, dsi, dd) = build(jgcm, 21, 32, 5, 600)
    step = lambda: jgcm.Atmosphere_Update(dev, datmo, dvc, dsi, dd, HS, carry_gradients=True,
                                          rotate=True)
    step()
    jgcm.Update_Init_Step(dsi)
    step()
    names = {f: getattr(dd, f).data_ptr() for f in ("spe_vor_p", "spe_vor_c", "spe_vor_n",
                                                    "grid_t_c")}
    for repeat in (1, 2, 4):
        with pytest.raises(AssertionError, match="multiple of 3"):
            jgcm.Step_Graph(dev, step, nfields=8 * 5 + 4, repeat=repeat)
        assert {f: getattr(dd, f).data_ptr() for f in names} == names
    g = jgcm.Step_Graph(dev, step, nfields=8 * 5 + 4, repeat=3)
    g.launch()
    step()                                  # an eager rotated step renames the state
    with pytest.raises(AssertionError, match="rotated since the capture"):
        g.launch()
    step()
    step()                                  # three renamings later the names are back
    g.launch()
    dev.synchronize()
    assert bool(torch.isfinite(dd.grid_t_c).all())
    g.close()


SW_FIELDS = [f"{f}_{s}" for f in ("spe_vor", "spe_div", "spe_lnps", "grid_u", "grid_v", "grid_ps")
             for s in ("p", "c")] + ["grid_vor", "grid_div"]


@pytest.mark.parametrize("concurrent", [False, True])
def test_rotated_shallow_water_ensemble_equals_copying(jgcm, concurrent):
    """The mixed ensemble with rotated steps (eager, then graphs of three
    steps) leaves p and c of every group equal, bit for bit, to the copying
    steps (eager, then one-step graphs)."""
    from jgcm_amd.ensemble import SW_Ensemble
    groups = [(21, 32, 3, 600), (42, 64, 2, 600), (85, 128, 2, 300)]
    A = SW_Ensemble(groups, seed=5, precision="fp64", concurrent=concurrent)
    B = SW_Ensemble(groups, seed=5, precision="fp64", concurrent=concurrent, rotate=True)
    for E in (A, B):
        E.step()
        E.step()
    A.capture()
    B.capture()
    for _ in range(6):
        A.launch()
    for _ in range(2):
        B.launch()
    A.synchronize()
    B.synchronize()
    for ga, gb in zip(A.groups, B.groups):
        assert ga.steps == gb.steps == 8
        for f in SW_FIELDS:
            assert torch.equal(getattr(ga.d, f), getattr(gb.d, f)), (f"T{ga.N}", f)
    A.graph.close()
    B.graph.close()


def test_rotated_galewsky_days_equal_copying(jgcm):
    """The Galewsky driver with rotated steps (3-step graphs) and with the
    copying step (1-step graphs) give the same state bit for bit after a day."""
    from jgcm_amd.experiments import Galewsky_Experiment
    out = []
    for rotate in (False, True):
        e = Galewsky_Experiment(N=42, nθ=64, dt=1200, precision="fp64", rotate=rotate)
        e.run_days(1)
        e.mesh.synchronize()
        out.append({f: getattr(e.d, f).clone() for f in SW_FIELDS})
        e.close()
    for f in SW_FIELDS:
        assert torch.equal(out[0][f], out[1][f]), f
