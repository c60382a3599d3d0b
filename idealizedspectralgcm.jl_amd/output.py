"""Output_Manager mirror: Held-Suarez diagnostics accumulated on the device.

Reference: src/Atmos_Spectral/Output_Manager.jl:4-145.  Update_Output! adds
the zonal means of T, T_eq, u, v and ps itself into per-day device
accumulators (one kernel per step, no host copy); Finalize_Output! copies the
accumulators once, divides by the per-day counts and forms the post-spin-up
means.  The reference's JLD2 ``@save`` becomes ``numpy.savez`` (JLD2 is a
Julia format); the plotting helpers are out of scope.  Array shapes follow
the reference read in C order: daily zonal means ``[n_day][nd][nθ]``, daily
ps ``[n_day][nθ][nλ]``.
"""
import ctypes

import numpy as np
import torch

from ._lib import call
from .mesh import _ptr


class Output_Manager:
    """``Output_Manager(mesh, vert_coord, start_time, end_time, spinup_day)``
    (Output_Manager.jl:41-79)."""

    def __init__(self, mesh, vert_coord, start_time, end_time, spinup_day):
        self.mesh = mesh
        self.nλ, self.nθ, self.nd = mesh.nλ, mesh.nθ, mesh.nd
        self.day_to_sec = 86400
        self.start_time, self.end_time = int(start_time), int(end_time)
        self.current_time = int(start_time)
        self.spinup_day = int(spinup_day)
        self.λc, self.θc = mesh.λc, mesh.θc
        bk = vert_coord.bk
        self.σc = (bk[1:] + bk[:-1]) / 2.0
        self.n_day = int((end_time - start_time) / self.day_to_sec)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device="cuda")
        n, nd, nθ, nλ = self.n_day, self.nd, self.nθ, self.nλ
        self._t, self._teq = z(n, nd, nθ), z(n, nd, nθ)
        self._u, self._v = z(n, nd, nθ), z(n, nd, nθ)
        self._ps = z(n, nθ, nλ)
        self.n_daily_mean = np.zeros(n)
        self.finalized = False


def Update_Output(output_manager, dyn_data, current_time):
    """Output_Manager.jl:81-107."""
    om = output_manager
    assert current_time > om.current_time
    om.current_time = current_time
    i_day = (current_time - om.start_time - 1) // om.day_to_sec   # 0-based day slot
    if i_day + 1 > om.n_day:
        print("Warning: i_day > n_day in Output_Manager:Update!")
        return
    d = dyn_data
    call("jgcm_zonal_mean_accumulate", om.mesh.handle, _ptr(d.grid_t_c), _ptr(d.grid_t_eq),
         _ptr(d.grid_u_c), _ptr(d.grid_v_c), _ptr(d.grid_ps_c), om.nd, _ptr(om._t[i_day]),
         _ptr(om._teq[i_day]), _ptr(om._u[i_day]), _ptr(om._v[i_day]), _ptr(om._ps[i_day]))
    om.n_daily_mean[i_day] += 1


def Finalize_Output(output_manager, save_file_name="None", mean_save_file_name="None"):
    """Output_Manager.jl:109-145."""
    om = output_manager
    om.mesh.synchronize()
    cnt = om.n_daily_mean.copy()
    div = np.where(cnt > 0, cnt, np.nan)
    om.t_daily_zonal_mean = om._t.cpu().numpy() / div[:, None, None]
    om.t_eq_daily_zonal_mean = om._teq.cpu().numpy() / div[:, None, None]
    om.u_daily_zonal_mean = om._u.cpu().numpy() / div[:, None, None]
    om.v_daily_zonal_mean = om._v.cpu().numpy() / div[:, None, None]
    om.ps_daily_mean = om._ps.cpu().numpy() / div[:, None, None]
    om.n_daily_mean[:] = 1.0
    s = slice(om.spinup_day, om.n_day)
    om.t_zonal_mean = om.t_daily_zonal_mean[s].mean(axis=0)
    om.t_eq_zonal_mean = om.t_eq_daily_zonal_mean[s].mean(axis=0)
    om.u_zonal_mean = om.u_daily_zonal_mean[s].mean(axis=0)
    om.v_zonal_mean = om.v_daily_zonal_mean[s].mean(axis=0)
    om.ps_mean = om.ps_daily_mean[s].mean(axis=0)
    om.finalized = True
    if save_file_name != "None":
        np.savez(save_file_name, t_daily_zonal_mean=om.t_daily_zonal_mean,
                 t_eq_daily_zonal_mean=om.t_eq_daily_zonal_mean,
                 u_daily_zonal_mean=om.u_daily_zonal_mean, v_daily_zonal_mean=om.v_daily_zonal_mean,
                 ps_daily_mean=om.ps_daily_mean, σc=om.σc, θc=np.asarray(om.θc))
    if mean_save_file_name != "None":
        np.savez(mean_save_file_name, t_zonal_mean=om.t_zonal_mean, u_zonal_mean=om.u_zonal_mean,
                 v_zonal_mean=om.v_zonal_mean)


def Zonal_Mean(grid_dat):
    """Output_Manager.jl:276-280, on a device grid field -> host [nd][nθ]."""
    return grid_dat.mean(dim=-1).cpu().numpy()
