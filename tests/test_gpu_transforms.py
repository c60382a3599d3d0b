"""GPU parity of the transform core against the FP64 oracle (through the C ABI).

Tolerances: FP64 path 1e-12 relative (max-abs error / max-abs value); the
FP32 and split-FP32 Legendre variants 1e-6 relative per level-field
(2-norm), the accuracy required of the reduced-precision transforms.
"""
import numpy as np
import pytest
import torch

from oracle import SpectralSphericalMesh

pytestmark = pytest.mark.gpu

RADIUS = 6371.0e3
# 128 and 256 longitudes: the four-step register FFT; 32, 64 and 512: the
# radix-4 LDS FFT; 96 and 192: the direct-DFT path (not powers of two)
TRUNCS = [(21, 32), (42, 64), (85, 128),
          (31, 48), (63, 96), (10, 16)]
_CACHE = {}


def pair(jgcm, N, nlat, nd, precision="fp64", radius=RADIUS):
    key = (N, nlat, nd, radius)
    if key not in _CACHE:
        _CACHE[key] = (jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nd, radius),
                       SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, nd, radius))
    dev, ref = _CACHE[key]
    dev.set_precision(precision)
    return dev, ref


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def field_rel(a, b):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return (np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max()


def random_spec(ref, nlev, rng, top=True):
    """Valid spectral coefficients: m=0 real, lower triangle zero."""
    N, ns = ref.num_fourier, ref.num_spherical
    s = rng.standard_normal((nlev, ns + 1, N + 1)) + 1j * rng.standard_normal((nlev, ns + 1, N + 1))
    s[:, :, 0] = s[:, :, 0].real
    for m in range(N + 1):
        s[:, :m, m] = 0
    if not top:
        s[:, ns, :] = 0
    return s


def cuda(x):
    return torch.tensor(np.ascontiguousarray(x), device="cuda")


@pytest.mark.parametrize("N,nlat", TRUNCS)
def test_forward_inverse_fp64(jgcm, N, nlat):
    nd = 5
    dev, ref = pair(jgcm, N, nlat, nd)
    rng = np.random.default_rng(N)
    g = rng.standard_normal((nd, nlat, 2 * nlat))
    s_ref = ref.trans_grid_to_spherical(g, ref.spec_zeros())
    sd = dev.spectral_zeros()
    jgcm.Trans_Grid_To_Spherical(dev, cuda(g), sd)
    assert rel(sd.cpu().numpy(), s_ref) < 1e-12
    sp = random_spec(ref, nd, rng)
    p_ref = ref.trans_spherical_to_grid(sp, ref.grid_zeros())
    pd = dev.grid_zeros()
    jgcm.Trans_Spherical_To_Grid(dev, cuda(sp), pd)
    assert rel(pd.cpu().numpy(), p_ref) < 1e-12


@pytest.mark.parametrize("N,nlat", TRUNCS + [(127, 256)])
def test_batched_61_fields_fp64(jgcm, N, nlat):
    """61 level-fields in one launch (vor 20, div 20, T 20, lnps 1)."""
    B = 61
    dev, ref = pair(jgcm, N, nlat, B)
    rng = np.random.default_rng(100 + N)
    sp = random_spec(ref, B, rng, top=False)
    p_ref = ref.trans_spherical_to_grid(sp, ref.grid_zeros())
    pd = dev.grid_zeros()
    jgcm.Trans_Spherical_To_Grid(dev, cuda(sp), pd)
    assert rel(pd.cpu().numpy(), p_ref) < 1e-12
    sd = dev.spectral_zeros()
    jgcm.Trans_Grid_To_Spherical(dev, pd, sd)
    # band-limited round trip reproduces the coefficients (degree <= N)
    assert rel(sd.cpu().numpy(), sp) < 1e-11


@pytest.mark.parametrize("precision", ["fp32", "split32", "bf16x3"])
@pytest.mark.parametrize("N,nlat", TRUNCS)
def test_reduced_precision_legendre(jgcm, N, nlat, precision):
    B = 61
    dev, ref = pair(jgcm, N, nlat, B, precision)
    try:
        rng = np.random.default_rng(7 + N)
        sp = random_spec(ref, B, rng, top=False)
        p_ref = ref.trans_spherical_to_grid(sp, ref.grid_zeros())
        pd = dev.grid_zeros()
        jgcm.Trans_Spherical_To_Grid(dev, cuda(sp), pd)
        e_inv = field_rel(pd.cpu().numpy(), p_ref)
        sd = dev.spectral_zeros()
        jgcm.Trans_Grid_To_Spherical(dev, cuda(p_ref), sd)
        e_fwd = field_rel(sd.cpu().numpy(), sp)
        assert e_inv < 1e-6 and e_fwd < 1e-6, (e_inv, e_fwd)
        # a reduced-precision path really ran (FP64 would sit near 1e-15)
        assert e_inv > 1e-11 and e_fwd > 1e-11, (e_inv, e_fwd)
        if precision in ("split32", "bf16x3"):
            assert e_inv < 2e-7 and e_fwd < 2e-7, (e_inv, e_fwd)
    finally:
        dev.set_precision("fp64")


def test_forward_leaves_lower_triangle(jgcm):
    dev, ref = pair(jgcm, 21, 32, 1)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((1, 32, 64))
    init = np.full((1, 23, 22), 7.0 + 7.0j)
    s_ref = ref.trans_grid_to_spherical(g, init.copy())
    sd = cuda(init)
    jgcm.Trans_Grid_To_Spherical(dev, cuda(g), sd)
    out = sd.cpu().numpy()
    assert rel(out, s_ref) < 1e-12
    assert np.all(out[0, 22, :] == 0)


@pytest.mark.parametrize("N,nlat", TRUNCS)
def test_vor_div_uv(jgcm, N, nlat):
    nd = 3
    dev, ref = pair(jgcm, N, nlat, nd)
    rng = np.random.default_rng(11 + N)
    u, v = rng.standard_normal((2, nd, nlat, 2 * nlat))
    vr, dr = ref.vor_div_from_grid_uv(u, v, ref.spec_zeros(), ref.spec_zeros())
    vd, dd = dev.spectral_zeros(), dev.spectral_zeros()
    jgcm.Vor_Div_From_Grid_UV(dev, cuda(u), cuda(v), vd, dd)
    assert rel(vd.cpu().numpy(), vr) < 1e-12
    assert rel(dd.cpu().numpy(), dr) < 1e-12
    vs, ds = random_spec(ref, nd, rng), random_spec(ref, nd, rng)
    ur, vvr = ref.uv_grid_from_vor_div(vs, ds, ref.grid_zeros(), ref.grid_zeros())
    ud, vvd = dev.grid_zeros(), dev.grid_zeros()
    jgcm.UV_Grid_From_Vor_Div(dev, cuda(vs), cuda(ds), ud, vvd)
    assert rel(ud.cpu().numpy(), ur) < 1e-12
    assert rel(vvd.cpu().numpy(), vvr) < 1e-12
    uc, vc = dev.spectral_zeros(), dev.spectral_zeros()
    ucr, vcr = ref.compute_ucos_vcos_from_vor_div(vs, ds, ref.spec_zeros(), ref.spec_zeros())
    jgcm.Compute_Ucos_Vcos_From_Vor_Div(dev, cuda(vs), cuda(ds), uc, vc)
    assert rel(uc.cpu().numpy(), ucr) < 1e-14
    assert rel(vc.cpu().numpy(), vcr) < 1e-14


@pytest.mark.parametrize("N,nlat", TRUNCS[:2] + TRUNCS[3:4])
def test_gradients_advection_laplacian(jgcm, N, nlat):
    nd = 2
    dev, ref = pair(jgcm, N, nlat, nd)
    rng = np.random.default_rng(5 + N)
    h = random_spec(ref, nd, rng)
    a, b = ref.compute_gradient_cos(h, ref.spec_zeros(), ref.spec_zeros())
    ad, bd = dev.spectral_zeros(), dev.spectral_zeros()
    jgcm.Compute_Gradient_Cos(dev, cuda(h), ad, bd)
    assert rel(ad.cpu().numpy(), a) < 1e-14 and rel(bd.cpu().numpy(), b) < 1e-14
    gl, gt = ref.compute_gradients(h, ref.grid_zeros(), ref.grid_zeros())
    gld, gtd = dev.grid_zeros(), dev.grid_zeros()
    jgcm.Compute_Gradients(dev, cuda(h), gld, gtd)
    assert rel(gld.cpu().numpy(), gl) < 1e-12 and rel(gtd.cpu().numpy(), gt) < 1e-12
    u, v, dh = rng.standard_normal((3, nd, nlat, 2 * nlat))
    dh_ref = ref.add_horizontal_advection(h, u, v, dh.copy())
    dhd = cuda(dh)
    jgcm.Add_Horizontal_Advection(dev, cuda(h), cuda(u), cuda(v), dhd)
    assert rel(dhd.cpu().numpy(), dh_ref) < 1e-12
    lap = ref.apply_laplacian(h.copy())
    hd = cuda(h)
    jgcm.Apply_Laplacian(dev, hd)
    assert rel(hd.cpu().numpy(), lap) < 1e-15
    g = rng.standard_normal((nd, nlat, 2 * nlat))
    m_ref = ref.area_weighted_global_mean(g)
    assert abs(jgcm.Area_Weighted_Global_Mean(dev, cuda(g)) - m_ref) < 1e-14 * max(1, abs(m_ref)) * 10
    gc = dev.grid_zeros()
    jgcm.Multiply_By_Cos(dev, cuda(g), gc)
    assert rel(gc.cpu().numpy(), ref.multiply_by_cos(g, np.empty_like(g))) < 1e-15
    jgcm.Divide_By_Cos(dev, gc)
    assert rel(gc.cpu().numpy(), g) < 1e-14


@pytest.mark.parametrize("N,nlat", TRUNCS[:2] + TRUNCS[3:4])
def test_fourier_and_alpha(jgcm, N, nlat):
    nd = 2
    dev, ref = pair(jgcm, N, nlat, nd)
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal((2, nd, nlat, 2 * nlat))
    fa = ref.trans_grid_to_fourier(a, np.empty(a.shape, np.complex128))
    fb = ref.trans_grid_to_fourier(b, np.empty(b.shape, np.complex128))
    fad = torch.zeros(a.shape, dtype=torch.complex128, device="cuda")
    fbd = torch.zeros_like(fad)
    jgcm.Trans_Grid_To_Fourier(dev, cuda(a), fad)
    jgcm.Trans_Grid_To_Fourier(dev, cuda(b), fbd)
    assert rel(fad.cpu().numpy(), fa) < 1e-13
    for isign in (-1.0, 1.0):
        al = ref.compute_alpha_operator(fa, fb, isign, ref.spec_zeros())
        ald = dev.spectral_zeros()
        jgcm.Compute_Alpha_Operator(dev, fad, fbd, isign, ald)
        assert rel(ald.cpu().numpy(), al) < 1e-12


def test_shape_assertions(jgcm):
    dev, _ = pair(jgcm, 21, 32, 4)
    with pytest.raises(AssertionError):
        jgcm.Trans_Grid_To_Spherical(dev, dev.grid_zeros(2), dev.spectral_zeros(2))
    with pytest.raises(AssertionError):
        jgcm.Trans_Grid_To_Spherical(dev, dev.grid_zeros(4), dev.spectral_zeros(1))
    with pytest.raises(AssertionError):
        jgcm.Spectral_Spherical_Mesh(21, 22, 31, 64, 1, 1.0)   # odd nθ


def test_mesh_tables_match_oracle(jgcm):
    dev, ref = pair(jgcm, 42, 64, 1)
    for name, rname in [("sinθ", "sin_theta"), ("wts", "wts"), ("cosθ", "cos_theta"),
                        ("qnm", "qnm"), ("dqnm", "dqnm"), ("qwg", "qwg"), ("epsilon", "epsilon"),
                        ("laplacian_eig", "laplacian_eig"), ("coef_uvm", "coef_uvm"),
                        ("coef_uvp", "coef_uvp"), ("coef_dθm", "coef_dthm"),
                        ("coef_alp_a", "coef_alp_a"), ("coef_alp_b", "coef_alp_b"),
                        ("θe", "theta_e"), ("λc", "lam_c")]:
        a, b = getattr(dev, name), getattr(ref, rname)
        assert a.shape == b.shape, name
        assert rel(a, b) < 1e-13, name
