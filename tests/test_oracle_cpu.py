"""The C++ FP64 restatement of the transform pair (oracle/sht_cpu.cpp, the
CPU baseline of bench.py) against the numpy oracle, on the CPU: the same
transforms to rounding at T21, T42 and T31 on the 96 x 48 grid (the direct
DFT branch), on 1 and several threads, and the transforms' structure
(degree num_spherical zeroed, lower triangle untouched)."""
import numpy as np
import pytest

from oracle import SpectralSphericalMesh, cpu


@pytest.fixture(scope="module", autouse=True)
def built():
    try:
        cpu.load()
    except RuntimeError:
        cpu.build()


@pytest.mark.parametrize("N,nlat,nlon,nlev", [(21, 32, 64, 3), (42, 64, 128, 5), (31, 48, 96, 2)])
@pytest.mark.parametrize("threads", [1, 3])
def test_cpp_transforms_equal_numpy_oracle(N, nlat, nlon, nlev, threads):
    M = SpectralSphericalMesh(N, N + 1, nlat, nlon, nlev, 6.371e6)
    c = cpu.CpuSHT(M)
    rng = np.random.default_rng(N + threads)
    g = rng.standard_normal((nlev, nlat, nlon))
    s_ref = M.trans_grid_to_spherical(g, M.spec_zeros())
    s = np.full_like(s_ref, 7.0 + 7.0j)
    c.forward(g, s, threads)
    tri = np.triu(np.ones((N + 2, N + 1), bool), 1)    # [n, m] with n < m: untouched
    assert np.all(s[:, tri] == 7.0 + 7.0j)
    assert np.all(s[:, N + 1, :] == 0.0)
    s[:, tri] = 0.0
    s_ref[:, tri] = 0.0
    assert np.abs(s - s_ref).max() <= 1e-13 * np.abs(s_ref).max()
    # inverse of a spectrum with degree N+1 filled (the inverse sums it, SSM:196)
    spec = (rng.standard_normal(s_ref.shape) + 1j * rng.standard_normal(s_ref.shape))
    spec[:, tri] = 0.0
    spec[:, :, 0] = spec[:, :, 0].real
    p_ref = M.trans_spherical_to_grid(spec, M.grid_zeros())
    p = np.zeros_like(p_ref)
    c.inverse(spec, p, threads)
    assert np.abs(p - p_ref).max() <= 1e-13 * np.abs(p_ref).max()
