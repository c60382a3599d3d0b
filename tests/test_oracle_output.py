"""Pins for oracle.output (CPU): Output_Manager.jl day bucketing and means."""
import numpy as np

from oracle.output import OutputManager, zonal_mean


class _D:
    pass


def _state(nd, nlat, nlon, tval, m=3):
    d = _D()
    lam = np.linspace(0, 2 * np.pi, nlon + 1)[:nlon]
    wave = np.cos(m * lam)[None, None, :]
    d.grid_t_c = tval + 5 * wave * np.ones((nd, nlat, 1))
    d.grid_t_eq = np.full((nd, nlat, nlon), 250.0)
    d.grid_u_c = np.ones((nd, nlat, nlon)) * np.arange(nlat)[None, :, None]
    d.grid_v_c = 2.0 * wave * np.ones((nd, nlat, 1))
    d.grid_ps_c = np.full((1, nlat, nlon), 1e5)
    return d


def test_zonal_mean_removes_waves_and_keeps_means():
    d = _state(3, 4, 16, 280.0)
    np.testing.assert_allclose(zonal_mean(d.grid_t_c), 280.0, atol=1e-12)
    np.testing.assert_allclose(zonal_mean(d.grid_v_c), 0.0, atol=1e-14)


def test_day_bucketing_and_spinup_mean():
    nd, nlat, nlon = 2, 4, 8
    om = OutputManager(nlon, nlat, nd, np.linspace(0, 1, nd + 1), 0, 3 * 86400, 1)
    dt = 43200
    for k in range(1, 7):                        # two steps per day, days 1..3
        om.update(_state(nd, nlat, nlon, 270.0 + (k - 1) // 2), k * dt)
    om.update(_state(nd, nlat, nlon, 999.0), 7 * dt)   # beyond n_day: ignored
    np.testing.assert_array_equal(om.count, [2, 2, 2])
    om.finalize()
    np.testing.assert_allclose(om.t[:, 0, 0], [270.0, 271.0, 272.0])
    np.testing.assert_allclose(om.t_zonal_mean, 271.5)       # days 2..3
    np.testing.assert_allclose(om.u_zonal_mean[0], np.arange(nlat))
    np.testing.assert_allclose(om.ps_mean, 1e5)
