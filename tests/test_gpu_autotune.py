"""FFT workgroup shapes on the device: every shape the tuner may pick gives
bit-identical transforms (only the assignment of latitudes to workgroups
changes); out-of-range shapes are rejected; the tuner applies and stores its
choice, and a cached choice is applied without timing."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def at(jgcm):
    from jgcm_amd import autotune
    return autotune


@pytest.mark.parametrize("N,nlat,nf", [(42, 64, 20), (85, 128, 61)])
def test_shapes_give_identical_transforms(jgcm, at, N, nlat, nf):
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nf, 6.371e6)
    g = torch.Generator(device="cpu").manual_seed(3)
    grid = torch.randn((nf, nlat, 2 * nlat), generator=g, dtype=torch.float64).cuda()
    spec0, grid0 = mesh.spectral_zeros(nf), mesh.grid_zeros(nf)
    jgcm.Trans_Grid_To_Spherical(mesh, grid, spec0)
    jgcm.Trans_Spherical_To_Grid(mesh, spec0, grid0)
    mesh.synchronize()
    limit = at.get_fft_shape(mesh)[2]
    assert limit >= 4
    for c in [c for c in at.CANDIDATES if c <= limit] + [limit, 3]:
        at.set_fft_shape(mesh, c, c)
        spec, out = mesh.spectral_zeros(nf), mesh.grid_zeros(nf)
        jgcm.Trans_Grid_To_Spherical(mesh, grid, spec)
        jgcm.Trans_Spherical_To_Grid(mesh, spec, out)
        mesh.synchronize()
        assert torch.equal(spec, spec0), c
        assert torch.equal(out, grid0), c
    with pytest.raises(AssertionError):   # argument errors, like the reference @assert
        at.set_fft_shape(mesh, limit + 1, 0)
    with pytest.raises(AssertionError):   # argument errors, like the reference @assert
        at.set_fft_shape(mesh, 0, -1)
    at.set_fft_shape(mesh, 0, 0)
    assert at.get_fft_shape(mesh)[:2] == (0, 0)


@pytest.mark.parametrize("prec", ["fp64", "split32", "bf16x3", "fp32"])
@pytest.mark.parametrize("N,nlat,nf", [(21, 32, 7), (85, 128, 61)])
def test_legendre_tiles_give_identical_transforms(jgcm, at, prec, N, nlat, nf):
    """Every Legendre tile shape gives bit-identical transforms (each
    coefficient's sum runs over the same latitudes in the same order; the
    tile count only changes which workgroup computes which rows), for plain
    transforms and for the operator paths (derivative tables, batched
    segments) that stage coefficients differently."""
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nf, 6.371e6, precision=prec)
    g = torch.Generator(device="cpu").manual_seed(4)
    grid = torch.randn((nf, nlat, 2 * nlat), generator=g, dtype=torch.float64).cuda()
    grid2 = torch.randn((nf, nlat, 2 * nlat), generator=g, dtype=torch.float64).cuda()

    def run():
        spec = mesh.spectral_zeros(nf)
        out = mesh.grid_zeros(nf)
        jgcm.Trans_Grid_To_Spherical(mesh, grid, spec)
        jgcm.Trans_Spherical_To_Grid(mesh, spec, out)
        vor, div = mesh.spectral_zeros(nf), mesh.spectral_zeros(nf)
        jgcm.Vor_Div_From_Grid_UV(mesh, grid, grid2, vor, div)
        u, v = mesh.grid_zeros(nf), mesh.grid_zeros(nf)
        jgcm.UV_Grid_From_Vor_Div(mesh, vor, div, u, v)
        mesh.synchronize()
        return [spec, out, vor, div, u, v]

    assert at.get_legendre_shape(mesh) == (0, 0)
    base = run()
    for ft in at.TILE_CANDIDATES:
        for it in at.TILE_CANDIDATES:
            at.set_legendre_shape(mesh, ft, it)
            assert at.get_legendre_shape(mesh) == (ft, it)
            for a, b in zip(run(), base):
                assert torch.equal(a, b), (ft, it)
    for bad in [(3, 0), (0, 8), (-1, 1)]:
        with pytest.raises(AssertionError):   # argument errors, like the reference @assert
            at.set_legendre_shape(mesh, *bad)
    at.set_legendre_shape(mesh, 0, 0)


def test_tune_stores_and_cache_applies(jgcm, at, tmp_path):
    mesh = jgcm.Spectral_Spherical_Mesh(42, 43, 64, 128, 20, 6.371e6)
    cache = at.Tuning_Cache(str(tmp_path / "tuning.json"))
    e = at.Tune_Shapes(mesh, 20, reps=5, cache=cache)
    assert at.get_fft_shape(mesh)[:2] == (e["fold_ppw"], e["unfold_ppw"])
    assert at.get_legendre_shape(mesh) == (e["fwd_tiles"], e["inv_tiles"])
    assert e["fwd_tiles"] in at.TILE_CANDIDATES and e["inv_tiles"] in at.TILE_CANDIDATES
    assert set(e["forward_us"]) == set(e["inverse_us"]) and len(e["forward_us"]) >= 4
    assert set(e["legendre_forward_us"]) == {"1", "2", "4"}
    other = jgcm.Spectral_Spherical_Mesh(42, 43, 64, 128, 20, 6.371e6)
    again = at.Apply_Tuning(other, 20, cache=at.Tuning_Cache(str(tmp_path / "tuning.json")))
    assert again["fold_ppw"] == e["fold_ppw"] and again["unfold_ppw"] == e["unfold_ppw"]
    assert at.get_fft_shape(other)[:2] == (e["fold_ppw"], e["unfold_ppw"])
    assert at.get_legendre_shape(other) == (e["fwd_tiles"], e["inv_tiles"])
    assert at.Apply_Tuning(other, 21, cache=cache) is None     # no entry for 21 fields
