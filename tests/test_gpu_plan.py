"""Transform plans (jgcm_plan_*): the batched transforms of several meshes of
mixed truncation, with the Legendre stage of all of them in one work-list
launch, equal the one-mesh batched calls bit for bit (every tile runs the
arithmetic of the one-mesh launch), whatever the number of persistent
workgroups; the work list holds every tile of the one-mesh launches exactly
once; and the persistent form of the mixed shallow-water ensemble steps every
instance exactly like the other forms and like the instance run alone."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RADIUS = 6371.0e3
SHAPES = [(21, 32), (42, 64), (85, 128)]


def random_spec(N, nlev, rng):
    s = rng.standard_normal((nlev, N + 2, N + 1)) + 1j * rng.standard_normal((nlev, N + 2, N + 1))
    s[:, :, 0] = s[:, :, 0].real
    for m in range(N + 1):
        s[:, :m, m] = 0
    return torch.tensor(s, device="cuda")


def shared_meshes(jgcm, prec, levels, tiles=None):
    from jgcm_amd._lib import call
    meshes = []
    for k, (N, nlat) in enumerate(SHAPES):
        m = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, levels[k], RADIUS, precision=prec)
        if meshes:
            call("jgcm_mesh_share_stream", m.handle, meshes[0].handle)
        if tiles is not None:
            call("jgcm_mesh_set_legendre_shape", m.handle, tiles[k][0], tiles[k][1])
        meshes.append(m)
    return meshes


def inverse_batches(jgcm, meshes, rng, levels):
    """Per mesh: plain, ÷cos, u cos / v cos from (vor, div) and both
    gradients, with level counts that make 32-field chunks straddle arrays."""
    out = []
    for m, n in zip(meshes, levels):
        N = m.num_fourier
        vor, div, h = (random_spec(N, n, rng) for _ in range(3))
        snms = [vor, div, h, vor, vor, h, h]
        grids = [m.grid_zeros(n) for _ in snms]
        flags = [0, jgcm.BATCH_DIVCOS, 0, jgcm.BATCH_DIVCOS | jgcm.INV_UCOS,
                 jgcm.BATCH_DIVCOS | jgcm.INV_VCOS, jgcm.BATCH_DIVCOS | jgcm.INV_GRAD_LAMBDA,
                 jgcm.BATCH_DIVCOS | jgcm.INV_GRAD_THETA]
        snms_b = [None, None, None, div, div, None, None]
        out.append((m, snms, grids, flags, snms_b))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32", "split32", "bf16x3"])
@pytest.mark.parametrize("workgroups", [-1, 0, 7])
def test_inverse_plan_equals_one_mesh_batches(jgcm, prec, workgroups):
    from jgcm_amd.plan import INVERSE, Transform_Plan
    levels = [9, 5, 6]
    meshes = shared_meshes(jgcm, prec, levels, tiles=[(1, 1), (2, 2), (2, 4)])
    rng = np.random.default_rng(3)
    batches = inverse_batches(jgcm, meshes, rng, levels)
    plan = Transform_Plan(INVERSE, batches, workgroups=workgroups)
    plan.execute()
    meshes[0].synchronize()
    for m, snms, grids, flags, snms_b in batches:
        ref = [m.grid_zeros(g.shape[0]) for g in grids]
        jgcm.Trans_Spherical_To_Grid_Batch(m, snms, ref, flags, snms_b=snms_b)
        m.synchronize()
        for k, (a, b) in enumerate(zip(grids, ref)):
            assert torch.equal(a, b), (f"T{m.num_fourier}", k)
            assert float(b.abs().max()) > 0
    n, w = plan.info()
    assert n > 0 and w == {0: n, 7: 7}.get(workgroups, w) and 0 < w <= n
    plan.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "split32", "bf16x3"])
@pytest.mark.parametrize("workgroups", [-1, 0, 5])
def test_forward_plan_equals_one_mesh_batches(jgcm, prec, workgroups):
    from jgcm_amd.plan import FORWARD, Transform_Plan
    levels = [40, 3, 7]
    meshes = shared_meshes(jgcm, prec, levels, tiles=[(4, 1), (1, 2), (2, 2)])
    gen = torch.Generator(device="cuda").manual_seed(5)
    batches = []
    for m, n in zip(meshes, levels):
        grids = [torch.randn((n, m.nθ, m.nλ), dtype=torch.float64, device="cuda", generator=gen)
                 for _ in range(4)]
        snms = [m.spectral_zeros(n) for _ in range(4)]
        batches.append((m, grids, snms, [jgcm.BATCH_VORDIV, 0, 0, 0]))
    plan = Transform_Plan(FORWARD, batches, workgroups=workgroups)
    plan.execute()
    meshes[0].synchronize()
    for m, grids, snms, flags in batches:
        ref = [m.spectral_zeros(s.shape[0]) for s in snms]
        jgcm.Trans_Grid_To_Spherical_Batch(m, grids, ref, flags)
        m.synchronize()
        for k, (a, b) in enumerate(zip(snms, ref)):
            assert torch.equal(a, b), (f"T{m.num_fourier}", k)
            assert float(b.abs().max()) > 0
    plan.close()


def test_work_lists_hold_every_tile_once(jgcm):
    """Inverse: every (mesh, segment, field chunk, latitude group) of the
    one-mesh grids; forward: every (mesh, segment, row group, job chunk) that
    has rows; the workgroup ranges cover the list."""
    from jgcm_amd.plan import FORWARD, INVERSE, Transform_Plan
    levels = [33, 70, 12]
    tiles = [(1, 2), (2, 1), (4, 4)]
    meshes = shared_meshes(jgcm, "fp32", levels, tiles=tiles)
    rng = np.random.default_rng(1)
    inv = Transform_Plan(INVERSE, inverse_batches(jgcm, meshes, rng, levels), workgroups=13)
    items, starts = inv.work_list()
    expect = set()
    for g, (m, n) in enumerate(zip(meshes, levels)):
        B = 7 * n
        trows = 16 * tiles[g][1]
        for x in range(2 * (m.num_fourier + 1)):
            for y in range((B + 31) // 32):
                for z in range((m.nθ // 2 + trows - 1) // trows):
                    expect.add((g, x, y, z))
    got = [tuple(int(v) for v in it) for it in items]
    assert len(got) == len(set(got)) == len(expect) and set(got) == expect
    assert starts[0] == 0 and starts[-1] == len(got) and len(starts) == 14
    fwd_batches = []
    for m, n in zip(meshes, levels):
        grids = [m.grid_zeros(n) for _ in range(4)]
        fwd_batches.append((m, grids, [m.spectral_zeros(n) for _ in range(4)],
                            [jgcm.BATCH_VORDIV, 0, 0, 0]))
    fwd = Transform_Plan(FORWARD, fwd_batches, workgroups=0)
    items, starts = fwd.work_list()
    expect = set()
    for g, (m, n) in enumerate(zip(meshes, levels)):
        N, trows = m.num_fourier, 16 * tiles[g][0]
        # jobs: vor (n fields), div (n fields), one plain job of 2n fields
        chunks = 2 * ((n + 31) // 32) + (2 * n + 31) // 32
        for x in range(2 * (N + 1)):
            nfirst = x // 2 + x % 2
            nr = (N - nfirst) // 2 + 1 if nfirst <= N else 0
            for y in range((nr + trows - 1) // trows):
                for z in range(chunks):
                    expect.add((g, x, y, z))
    got = [tuple(int(v) for v in it) for it in items]
    assert len(got) == len(set(got)) == len(expect) and set(got) == expect
    assert list(starts) == list(range(len(got) + 1))
    inv.close()
    fwd.close()


def test_plan_refuses_mixed_streams_and_changed_precision(jgcm):
    from jgcm_amd._lib import JGCMError, call
    from jgcm_amd.plan import INVERSE, Transform_Plan
    a = jgcm.Spectral_Spherical_Mesh(21, 22, 32, 64, 2, RADIUS)
    b = jgcm.Spectral_Spherical_Mesh(42, 43, 64, 128, 2, RADIUS)
    rng = np.random.default_rng(0)
    batches = [(m, [random_spec(m.num_fourier, 2, rng)], [m.grid_zeros(2)]) for m in (a, b)]
    with pytest.raises(AssertionError, match="stream"):
        Transform_Plan(INVERSE, batches)
    call("jgcm_mesh_share_stream", b.handle, a.handle)
    with pytest.raises(AssertionError, match="twice"):
        Transform_Plan(INVERSE, [batches[0], batches[0]])
    plan = Transform_Plan(INVERSE, batches)
    plan.execute()
    call("jgcm_mesh_set_precision", b.handle, 1)
    with pytest.raises(JGCMError, match="precision"):
        plan.execute()
    call("jgcm_mesh_set_precision", b.handle, 0)
    plan.execute()
    a.synchronize()
    plan.close()


GROUPS = [(21, 32, 3, 600), (42, 64, 2, 600), (85, 128, 2, 300)]


@pytest.mark.parametrize("workgroups", [-1, 0])
def test_persistent_ensemble_equals_sequential_form(jgcm, workgroups):
    """Eager stage-by-stage steps and replays of the captured persistent step
    equal the sequential form (one Shallow_Water_Step per truncation) bit for
    bit, the Legendre stages of the three truncations running as one launch
    per direction."""
    from jgcm_amd.ensemble import SW_Ensemble
    A = SW_Ensemble(GROUPS, seed=2, precision="fp32")
    B = SW_Ensemble(GROUPS, seed=2, precision="fp32", persistent=True, workgroups=workgroups)
    C = SW_Ensemble(GROUPS, seed=2, precision="fp32", persistent=True, workgroups=workgroups)
    for E in (A, B, C):
        E.step()
    C.capture()
    for k in range(4):
        A.step()
        B.step()
        C.launch()
        for E in (A, B):
            E.synchronize()
        C.owner.synchronize()
        for ga, gb, gc in zip(A.groups, B.groups, C.groups):
            for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_v_c", "grid_ps_c"):
                a = getattr(ga.d, f)
                assert torch.equal(a, getattr(gb.d, f)), (k, f"T{ga.N}", f, "eager")
                assert torch.equal(a, getattr(gc.d, f)), (k, f"T{ga.N}", f, "graph")
    C.graph.close()


def test_persistent_ensemble_instances_equal_single_runs(jgcm):
    """Each instance of the persistent form equals the same instance run
    alone (a one-instance group, one-mesh calls) bit for bit."""
    from jgcm_amd.ensemble import SW_Ensemble, SW_Group
    E = SW_Ensemble(GROUPS, seed=9, precision="fp64", persistent=True)
    singles = []
    for g in E.groups:
        vor0 = g.d.spe_vor_c.cpu().numpy()
        singles.append([SW_Group(g.N, g.mesh.nθ, 1, g.dt, 0, precision="fp64",
                                 vor0=vor0[i:i + 1]) for i in range(g.count)])
    for _ in range(4):
        E.step()
        for row in singles:
            for s in row:
                s.step()
    E.synchronize()
    for g, row in zip(E.groups, singles):
        for i, s in enumerate(row):
            s.mesh.synchronize()
            for f in ("spe_vor_c", "spe_div_c", "spe_lnps_c", "grid_u_c", "grid_ps_c"):
                assert torch.equal(getattr(g.d, f)[i], getattr(s.d, f)[0]), (f"T{g.N}", i, f)
