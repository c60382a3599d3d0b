"""Adjoint transforms: dot-product tests <A x, y> = <x, A^T y> in FP64 and the
torch autograd wrappers (the rrule counterpart) against central differences.
A linear map's adjoint is pinned by the dot-product identity alone."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rspec(N, nlev, rng, upper_only=True):
    s = rng.standard_normal((nlev, N + 2, N + 1)) + 1j * rng.standard_normal((nlev, N + 2, N + 1))
    if upper_only:
        for m in range(N + 1):
            s[:, :m, m] = 0
    return s


def dot_spec(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


@pytest.mark.parametrize("N,nlat", [(21, 32), (42, 64), (85, 128)])
def test_dot_product_identities(jgcm, N, nlat):
    from jgcm_amd import adjoint
    nlev = 3
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nlev, 6.371e6)
    rng = np.random.default_rng(N)
    x = rng.standard_normal((nlev, nlat, 2 * nlat))
    s = rspec(N, nlev, rng)
    xd, sd = torch.tensor(x, device="cuda"), torch.tensor(s, device="cuda")
    # forward transform F: grid -> spectral
    Fx = mesh.spectral_zeros()
    jgcm.Trans_Grid_To_Spherical(mesh, xd, Fx)
    FTs = mesh.grid_zeros()
    adjoint.trans_grid_to_spherical_adjoint(mesh, sd, FTs)
    torch.cuda.synchronize()
    lhs = dot_spec(Fx.cpu().numpy(), s)
    rhs = float(np.sum(x * FTs.cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs)), (lhs, rhs)
    # inverse transform G: spectral -> grid (uses degrees m..N+1)
    Gs = mesh.grid_zeros()
    jgcm.Trans_Spherical_To_Grid(mesh, sd, Gs)
    GTx = mesh.spectral_zeros()
    adjoint.trans_spherical_to_grid_adjoint(mesh, xd, GTx)
    torch.cuda.synchronize()
    lhs = float(np.sum(Gs.cpu().numpy() * x))
    rhs = dot_spec(s, GTx.cpu().numpy())
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs)), (lhs, rhs)
    g = GTx.cpu().numpy()
    for m in range(N + 1):
        assert not g[:, :m, m].any()


@pytest.mark.parametrize("N,nlat", [(21, 32), (85, 128), (31, 48)])
def test_batched_adjoints_equal_single_calls(jgcm, N, nlat):
    """The batched adjoint transforms (one FFT and one Legendre launch for
    several arrays, ÷cos θ on the way in, accumulation into the targets) give
    the single calls' results bit for bit; 96 x 48 runs the direct-DFT path."""
    from jgcm_amd import adjoint
    nd = 3
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nd, 6.371e6)
    rng = np.random.default_rng(40 + N)

    def cuda(a):
        return torch.tensor(a, device="cuda")

    g = [cuda(rng.standard_normal((n, nlat, 2 * nlat))) for n in (nd, 1, nd)]
    divcos = [0, 1, 0]
    singles = []
    for x, dc in zip(g, divcos):
        xin = x.clone()
        if dc:
            jgcm.Divide_By_Cos(mesh, xin)
        out = mesh.spectral_zeros(x.shape[0])
        adjoint.trans_spherical_to_grid_adjoint(mesh, xin, out)
        singles.append(out)
    outs = [mesh.spectral_zeros(x.shape[0]).fill_(1.0) for x in g]   # overwritten, rest zeroed
    adjoint.trans_spherical_to_grid_adjoint_batch(mesh, g, outs, divcos=divcos)
    tgt = [cuda(rspec(N, x.shape[0], rng)) for x in g]
    acc = [t.clone() for t in tgt]
    adjoint.trans_spherical_to_grid_adjoint_batch(mesh, g, acc, divcos=divcos, accumulate=True)
    s = [cuda(rspec(N, n, rng)) for n in (nd, 1, nd)]
    fs = []
    for x in s:
        o = mesh.grid_zeros(x.shape[0])
        adjoint.trans_grid_to_spherical_adjoint(mesh, x, o)
        fs.append(o)
    fb = [mesh.grid_zeros(x.shape[0]) for x in s]
    adjoint.trans_grid_to_spherical_adjoint_batch(mesh, s, fb)
    mesh.synchronize()
    for k in range(3):
        assert torch.equal(outs[k], singles[k]), k
        assert torch.equal(acc[k], tgt[k] + singles[k]), k
        assert torch.equal(fb[k], fs[k]), k
    with pytest.raises(AssertionError):     # accumulated outputs must not overlap
        adjoint.trans_spherical_to_grid_adjoint_batch(mesh, [g[0], g[2]], [acc[0], acc[0]],
                                                      accumulate=True)


def test_autograd_matches_central_differences(jgcm):
    from jgcm_amd.adjoint import GridToSpherical, SphericalToGrid
    N, nlat, nlev = 21, 32, 2
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nlev, 6.371e6)
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.standard_normal((nlev, nlat, 2 * nlat)), device="cuda", requires_grad=True)
    w = torch.tensor(rspec(N, nlev, rng), device="cuda")

    def loss(xx):
        s = GridToSpherical.apply(mesh, xx)
        y = SphericalToGrid.apply(mesh, s * w)
        return (y * y).sum()

    L = loss(x)
    L.backward()
    g = x.grad.detach().cpu().numpy()
    for seed in range(4):
        d = np.random.default_rng(100 + seed).standard_normal(x.shape)
        d /= np.linalg.norm(d)
        dd = torch.tensor(d, device="cuda")
        h = 1e-3
        with torch.no_grad():
            fd = (loss(x + h * dd) - loss(x - h * dd)).item() / (2 * h)
        an = float(np.sum(g * d))
        assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (fd, an)


def test_add_batch_equals_axpby_and_rejects_overlap(jgcm):
    """jgcm_add_batch (the reverse steps' cotangent accumulation) adds each
    pair exactly like jgcm_axpby with unit coefficients, and refuses
    destinations that overlap."""
    import ctypes
    from jgcm_amd._lib import call
    mesh = jgcm.Spectral_Spherical_Mesh(21, 22, 32, 64, 1, 6.371e6)
    rng = np.random.default_rng(5)
    n = (1000, 37, 4096)
    dst = [torch.tensor(rng.standard_normal(k), device="cuda") for k in n]
    src = [torch.tensor(rng.standard_normal(k), device="cuda") for k in n]
    ref = [d.clone() for d in dst]
    P = ctypes.c_void_p
    for d, s in zip(ref, src):
        call("jgcm_axpby", mesh.handle, d.numel(), 1.0, P(d.data_ptr()), 1.0, P(s.data_ptr()),
             P(d.data_ptr()))
    call("jgcm_add_batch", mesh.handle, 3, (P * 3)(*[d.data_ptr() for d in dst]),
         (P * 3)(*[s.data_ptr() for s in src]), (ctypes.c_size_t * 3)(*n))
    mesh.synchronize()
    for a, b in zip(dst, ref):
        assert torch.equal(a, b)
    with pytest.raises(AssertionError, match="overlap"):
        call("jgcm_add_batch", mesh.handle, 2, (P * 2)(dst[2].data_ptr(), dst[2].data_ptr() + 800),
             (P * 2)(src[2].data_ptr(), src[1].data_ptr()), (ctypes.c_size_t * 2)(200, 37))
