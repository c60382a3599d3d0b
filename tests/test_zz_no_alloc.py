"""The allocation-free forward transform (jgcm_trans_grid_to_spherical_noalloc)
against the allocating one (jgcm_trans_grid_to_spherical), through the C ABI:
bit-identical results on two calls in a row, the second smaller, at T85 and
T42; sizes beyond the static scratch are refused before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

A = 6.371e6


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("N,nlat,sizes", [(85, 128, (61, 20)), (42, 64, (200, 3))])
def test_noalloc_forward_equals_allocating_forward(jgcm, N, nlat, sizes):
    lib = jgcm.lib()
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    rng = np.random.default_rng(N)
    for nlev in sizes:                       # the second call is smaller than the first
        g = torch.tensor(rng.standard_normal((nlev, nlat, 2 * nlat)), device="cuda")
        ref = mesh.spectral_zeros(nlev)
        out = torch.full_like(ref, complex(np.nan, np.nan))
        print(f"noalloc T{N} nlev={nlev}", flush=True)
        assert lib.jgcm_trans_grid_to_spherical(mesh.handle, _ptr(g), _ptr(ref), nlev) == 0
        rc = lib.jgcm_trans_grid_to_spherical_noalloc(mesh.handle, _ptr(g), _ptr(out), nlev)
        assert rc == 0, lib.jgcm_last_error()
        mesh.synchronize()
        # the lower triangle (n < m) is not written by either call
        tri = torch.ones(N + 2, N + 1, dtype=torch.bool, device="cuda").triu(1)
        out[:, tri] = ref[:, tri]
        assert torch.equal(out, ref)
        assert bool(torch.isfinite(ref.real).all())


def test_noalloc_refuses_more_than_its_scratch(jgcm):
    lib = jgcm.lib()
    N, nlat = 85, 128
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    g = mesh.grid_zeros(1)
    s = mesh.spectral_zeros(1)
    # 96 level-fields at T85 need 16.9 MB of staging; the pointers are never read
    rc = lib.jgcm_trans_grid_to_spherical_noalloc(mesh.handle, _ptr(g), _ptr(s), 96)
    assert rc == jgcm._lib.JGCM_ERR_ARG
    assert b"static scratch" in lib.jgcm_last_error()
    assert lib.jgcm_trans_grid_to_spherical_noalloc(mesh.handle, _ptr(g), _ptr(s), 1) == 0
    mesh.synchronize()
