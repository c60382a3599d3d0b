"""Batched transforms (jgcm_trans_*_batch): several arrays of ragged level
counts in one launch pair must give exactly the one-array results (the
per-field arithmetic is the same, so the comparison is bit-exact) and match
the FP64 oracle to 1e-12.  Segment sizes are chosen so that 32-field
Legendre chunks straddle array boundaries."""
import numpy as np
import pytest
import torch

from oracle import SpectralSphericalMesh

pytestmark = pytest.mark.gpu

RADIUS = 6371.0e3


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_spec(N, ns, nlev, rng):
    s = rng.standard_normal((nlev, ns + 1, N + 1)) + 1j * rng.standard_normal((nlev, ns + 1, N + 1))
    s[:, :, 0] = s[:, :, 0].real
    for m in range(N + 1):
        s[:, :m, m] = 0
    return s


def cuda(x):
    return torch.tensor(np.ascontiguousarray(x), device="cuda")


@pytest.mark.parametrize("N,nlat,prec", [(42, 64, "fp64"), (85, 128, "fp64"), (85, 128, "split32"),
                                         (85, 128, "bf16x3")])
def test_inverse_batch_equals_single_calls(jgcm, N, nlat, prec):
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 5, RADIUS, precision=prec)
    rng = np.random.default_rng(N)
    levels = [5, 1, 40, 3]
    flags = [0, jgcm.BATCH_DIVCOS, 0, jgcm.BATCH_DIVCOS]
    specs = [cuda(random_spec(N, N + 1, n, rng)) for n in levels]
    outs = [dev.grid_zeros(n) for n in levels]
    jgcm.Trans_Spherical_To_Grid_Batch(dev, specs, outs, flags)
    for s, o, f in zip(specs, outs, flags):
        single = dev.grid_zeros(s.shape[0])
        jgcm.Trans_Spherical_To_Grid(dev, s, single, batch=True)
        if f:
            jgcm.Divide_By_Cos(dev, single)
        torch.cuda.synchronize()
        # the batch divides by cos θ in the FFT epilogue, the single path in a
        # separate kernel: the same operation, so the same rounding
        assert torch.equal(o, single)
    if prec == "fp64":
        ref = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, RADIUS)
        for s, o, f in zip(specs, outs, flags):
            sn = s.cpu().numpy()
            g = np.concatenate([ref.trans_spherical_to_grid(sn[i:i + 1], ref.grid_zeros(1))
                                for i in range(sn.shape[0])])
            if f:
                g = g / ref.cos_theta[None, :, None]
            assert rel(o.cpu().numpy(), g) < 1e-12


@pytest.mark.parametrize("N,nlat,prec", [(42, 64, "fp64"), (85, 128, "fp64"), (85, 128, "split32"),
                                         (85, 128, "bf16x3")])
def test_forward_batch_equals_single_calls(jgcm, N, nlat, prec):
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 5, RADIUS, precision=prec)
    rng = np.random.default_rng(7 + N)
    levels = [5, 1, 33, 33, 2]
    flags = [0, 0, jgcm.BATCH_VORDIV, 0, 0]
    grids = [cuda(rng.standard_normal((n, nlat, 2 * nlat))) for n in levels]
    outs = [dev.spectral_zeros(n) for n in levels]
    jgcm.Trans_Grid_To_Spherical_Batch(dev, grids, outs, flags)
    singles = [dev.spectral_zeros(n) for n in levels]
    for k in (0, 1, 4):
        jgcm.Trans_Grid_To_Spherical(dev, grids[k], singles[k], batch=True)
    jgcm.Vor_Div_From_Grid_UV(dev, grids[2], grids[3], singles[2], singles[3], batch=True)
    torch.cuda.synchronize()
    for o, s in zip(outs, singles):
        assert torch.equal(o, s)
    if prec == "fp64":
        ref = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, 1, RADIUS)
        for k in (0, 1, 4):
            gk = grids[k].cpu().numpy()
            r = np.concatenate([ref.trans_grid_to_spherical(gk[i:i + 1], ref.spec_zeros(1))
                                for i in range(levels[k])])
            assert rel(outs[k].cpu().numpy(), r) < 1e-12
        gu, gv = grids[2].cpu().numpy(), grids[3].cpu().numpy()
        for i in range(0, 33, 8):
            vr, dr = ref.vor_div_from_grid_uv(gu[i:i + 1], gv[i:i + 1], ref.spec_zeros(1),
                                              ref.spec_zeros(1))
            assert rel(outs[2][i:i + 1].cpu().numpy(), vr) < 1e-12
            assert rel(outs[3][i:i + 1].cpu().numpy(), dr) < 1e-12


def test_batch_argument_errors(jgcm):
    dev = jgcm.Spectral_Spherical_Mesh(21, 22, 32, 64, 2, RADIUS)
    g = [dev.grid_zeros(2) for _ in range(5)]
    s = [dev.spectral_zeros(2) for _ in range(5)]
    V = jgcm.BATCH_VORDIV
    with pytest.raises(AssertionError):   # u without a following v
        jgcm.Trans_Grid_To_Spherical_Batch(dev, g[:2], s[:2], [0, V])
    with pytest.raises(AssertionError):   # 2 + 1 + 2 Legendre jobs > 4
        jgcm.Trans_Grid_To_Spherical_Batch(dev, g, s, [V, 0, 0, V, 0])
    with pytest.raises(AssertionError):   # unknown inverse operator
        jgcm.Trans_Spherical_To_Grid_Batch(dev, s[:1], g[:1], [16])
    with pytest.raises(AssertionError):   # u cos θ without the divergence
        jgcm.Trans_Spherical_To_Grid_Batch(dev, s[:1], g[:1], [jgcm.INV_UCOS])
    with pytest.raises(AssertionError):   # more than 12 arrays
        jgcm.Trans_Spherical_To_Grid_Batch(dev, s * 3, g * 3)


@pytest.mark.parametrize("prec", ["fp64", "split32", "bf16x3"])
def test_inverse_operators_equal_unfused(jgcm, prec):
    """Gradient and u cos / v cos operators applied while staging equal the
    stand-alone operator kernels followed by the plain transform and ÷cos θ."""
    N, nlat = 42, 64
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 3, RADIUS, precision=prec)
    rng = np.random.default_rng(3)
    h = cuda(random_spec(N, N + 1, 3, rng))
    vor = cuda(random_spec(N, N + 1, 3, rng))
    div = cuda(random_spec(N, N + 1, 3, rng))
    D = jgcm.BATCH_DIVCOS
    outs = [dev.grid_zeros(3) for _ in range(4)]
    jgcm.Trans_Spherical_To_Grid_Batch(
        dev, [h, h, vor, vor], outs,
        [D | jgcm.INV_GRAD_LAMBDA, D | jgcm.INV_GRAD_THETA, D | jgcm.INV_UCOS, D | jgcm.INV_VCOS],
        snms_b=[None, None, div, div])
    sl, st = dev.spectral_zeros(3), dev.spectral_zeros(3)
    jgcm.Compute_Gradient_Cos(dev, h, sl, st)
    su, sv = dev.spectral_zeros(3), dev.spectral_zeros(3)
    jgcm.Compute_Ucos_Vcos_From_Vor_Div(dev, vor, div, su, sv)
    for s, o in zip((sl, st, su, sv), outs):
        g = dev.grid_zeros(3)
        jgcm.Trans_Spherical_To_Grid(dev, s, g, batch=True)
        jgcm.Divide_By_Cos(dev, g)
        torch.cuda.synchronize()
        assert torch.equal(o, g)


def test_damped_leapfrog_batch_equals_single_calls(jgcm):
    N, nlat, nd = 42, 64, 4
    dev = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nd, RADIUS)
    rng = np.random.default_rng(9)
    for init in (True, False):
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, dev.laplacian_eig, 0.5, 600, True, 0, 86400)
        if not init:
            jgcm.Update_Init_Step(I)
        state = [[cuda(random_spec(N, N + 1, n, rng)) for _ in range(4)] for n in (nd, nd, 1)]
        a = [[t.clone() for t in f] for f in state]
        b = [[t.clone() for t in f] for f in state]
        damp = [True, False, True]
        jgcm.Damped_Leapfrog_Batch(I, [tuple(f) + (dm,) for f, dm in zip(a, damp)])
        for f, dm in zip(b, damp):
            if dm:
                jgcm.Compute_Spectral_Damping(I, f[2], f[1], f[0])
            jgcm.Filtered_Leapfrog_step(I, *f)
        torch.cuda.synchronize()
        for fa, fb in zip(a, b):
            for x, y in zip(fa, fb):
                assert torch.equal(x, y)
