"""Seeded inputs and initial-state formulas shared by the fixture generator
(make_golden.py, which runs the oracle) and the tests that check the fixtures
(test_golden.py on the CPU, test_gpu_golden.py through the HIP library).
Nothing here imports the oracle."""
import numpy as np

A, OMEGA = 6371000.0, 7.292e-5
HS = {"σ_b": 0.7, "k_f": 1.0, "k_a": 1.0 / 40.0, "k_s": 1.0 / 4.0, "ΔT_y": 60.0, "Δθ_z": 10.0}


def seeded_spec(N, nlev, seed):
    """Spectral state of real fields: lower triangle 0, m = 0 real, degree N+1 zero."""
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((nlev, N + 2, N + 1)) + 1j * rng.standard_normal((nlev, N + 2, N + 1))
    for m in range(N + 1):
        s[:, :m, m] = 0
    s[:, :, 0] = s[:, :, 0].real
    s[:, N + 1, :] = 0
    return s


def seeded_grid(nlev, nlat, seed):
    return np.random.default_rng(seed).standard_normal((nlev, nlat, 2 * nlat))


def rh4_grid_vorticity(sin_theta, cos_theta, lam):
    """Rossby-Haurwitz wave 4 vorticity (exp/Barotropic), [1][nθ][nλ]."""
    w = K = 7.848e-6
    mu = np.asarray(sin_theta)[:, None]
    c = np.asarray(cos_theta)[:, None]
    return np.ascontiguousarray((2 * w * mu - K * mu * c ** 4 * 30 * np.cos(4 * np.asarray(lam)[None, :]))[None])


def sw_forcing(theta, lam):
    """h_0 and the forced height h_eq of exp/Shallow_Water (SW.jl:57-77)."""
    th, lam = np.asarray(theta), np.asarray(lam)
    h_0 = 3.0e4
    h_amp, h_lon, h_lat, h_width, h_itcz, itcz_w = (2.0e4, np.pi / 2, 25 * np.pi / 180,
                                                    15 * np.pi / 180, 1.0e5, 4 * np.pi / 180)
    x2 = ((lam[None, :] - h_lon) / (2.0 * h_width)) ** 2
    y2 = ((th[:, None] - h_lat) / h_width) ** 2
    h_eq = (h_0 + h_amp * np.exp(-(x2 + y2)) + h_itcz * np.exp(-(th[:, None] / itcz_w) ** 2))[None]
    return h_0, np.ascontiguousarray(h_eq)


# seeds and sizes of the fixtures
T21 = dict(N=21, nlat=32, nlev=3, seed=2100, uv_seeds=(2110, 2111), vd_seeds=(2120, 2121),
           baro_steps=6, sw_steps=6, hs_nd=10, hs_dt=600, hs_steps=3, hs_seed=11)
T42 = dict(N=42, nlat=64, nlev=1, seed=4200)
T85 = dict(N=85, nlat=128, nlev=4, seed=8500, rows=(0, 1, 63, 64, 126, 127))
