"""The C++ FP64 restatements (the CPU baselines of bench.py) against the numpy
oracle, on the CPU:

* the transform pair (oracle/sht_cpu.cpp): the same transforms to rounding at
  T21, T42 and T31 on the 96 x 48 grid (the direct DFT branch), on 1 and
  several threads, and the transforms' structure (degree num_spherical
  zeroed, lower triangle untouched);
* the Held-Suarez step (oracle/hs_cpu.cpp): every Dyn_Data array a step
  writes equal to the numpy step's over an init step and three leapfrog
  steps, the same bits on 1 and several threads, and a 0.1 % change of one
  forcing constant (ΔT_y) visible far above the tolerance, so that the
  comparison resolves a term of the step."""
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


HS = {"σ_b": 0.7, "k_f": 1.0, "k_a": 1.0 / 40.0, "k_s": 1.0 / 4.0, "ΔT_y": 60.0, "Δθ_z": 10.0}


def hs_state(N=21, nlat=32, nd=10, dt=1200):
    """The seeded Held-Suarez initial state (even sigma, isothermal 300 K
    with 0.1 K noise) of the oracle, as in the GPU parity tests."""
    from oracle import FilteredLeapfrog
    from oracle import dynamics as odyn
    from oracle import primitive as P
    M = SpectralSphericalMesh(N, N + 1, nlat, 2 * nlat, nd, 6371000.0)
    vc = P.VertCoordinate(2 * nlat, nlat, nd, "even_sigma", "simmons_and_burridge",
                          "second_centered_wts", 1e5)
    atmo = odyn.AtmoData("hs", 2 * nlat, nlat, nd, True, True, False, False, M.sin_theta,
                         6371000.0, 7.292e-5)
    I = FilteredLeapfrog(0.04, 4, 1.15741e-4, M.laplacian_eig, 0.5, dt, True, 0, 86400)
    si = P.SemiImplicitSolver(vc, atmo, I, 1e5, np.full(nd, 300.0), M.wave_numbers)
    d = odyn.DynData("hs", N, N + 1, 2 * nlat, nlat, nd)
    noise = 0.1 * np.random.default_rng(5).standard_normal((nd, nlat, 2 * nlat))
    P.spectral_initialize_fields(M, atmo, vc, 1e5, 300.0, d.grid_geopots, d, grid_t_noise=noise)
    return M, atmo, vc, si, d


def hs_fields(d):
    # grid_d_full1 is the numpy step's scratch for one vertical advection
    return {a: getattr(d, a) for _, a in cpu.HS_STATE if a != "grid_d_full1"}


def worst_rel(a, b):
    return max(float(np.abs(a[k] - b[k]).max() / max(np.abs(a[k]).max(), 1e-300)) for k in a)


def run_steps(state, steps, stepper):
    for k in range(steps):
        stepper()
        if k == 0:
            state[3].update_init_step()


def test_cpp_held_suarez_step_equals_numpy_oracle():
    from oracle import primitive as P
    A, B, C = hs_state(), hs_state(), hs_state()
    hs_b = cpu.CpuHS(*B, HS)
    hs_c = cpu.CpuHS(*C, HS)
    run_steps(A, 4, lambda: P.atmosphere_update(*A, HS))
    run_steps(B, 4, lambda: hs_b.step(1))
    run_steps(C, 4, lambda: hs_c.step(3))
    fa, fb, fc = hs_fields(A[4]), hs_fields(B[4]), hs_fields(C[4])
    for k in fa:
        assert np.array_equal(fb[k], fc[k]), k           # thread count changes no bit
        assert np.all(np.isfinite(fb[k])), k
    # observed 7e-11 (the divergence is near zero in this state)
    assert worst_rel(fa, fb) <= 2e-9
    assert float(np.abs(fb["grid_u_c"]).max()) > 1e-3     # the state has moved


def test_cpp_held_suarez_comparison_resolves_the_forcing():
    from oracle import primitive as P
    A, B = hs_state(), hs_state()
    hs_b = cpu.CpuHS(*B, dict(HS, **{"ΔT_y": HS["ΔT_y"] * 1.001}))
    run_steps(A, 4, lambda: P.atmosphere_update(*A, HS))
    run_steps(B, 4, lambda: hs_b.step(2))
    assert worst_rel(hs_fields(A[4]), hs_fields(B[4])) > 1e-6
