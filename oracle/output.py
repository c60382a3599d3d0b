"""Output_Manager diagnostics (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates Output_Manager.jl:41-145 (Output_Manager ctor, Update_Output!,
Finalize_Output! without the file output) and Zonal_Mean (:276-280), with
arrays in C order: daily zonal means [n_day][nd][nθ], daily ps [n_day][nθ][nλ].
Pinned in tests/test_oracle_output.py (constant and single-wavenumber fields,
day bucketing, spin-up averaging).
"""
import numpy as np


def zonal_mean(grid):
    return grid.mean(axis=-1)


class OutputManager:
    def __init__(self, nlon, nlat, nd, bk, start_time, end_time, spinup_day):
        self.day_to_sec = 86400
        self.start_time, self.end_time, self.current_time = start_time, end_time, start_time
        self.spinup_day = spinup_day
        self.sigma_c = (bk[1:] + bk[:-1]) / 2.0
        self.n_day = int((end_time - start_time) / self.day_to_sec)
        n = self.n_day
        self.t = np.zeros((n, nd, nlat))
        self.teq = np.zeros((n, nd, nlat))
        self.u = np.zeros((n, nd, nlat))
        self.v = np.zeros((n, nd, nlat))
        self.ps = np.zeros((n, nlat, nlon))
        self.count = np.zeros(n)

    def update(self, d, current_time):
        assert current_time > self.current_time
        self.current_time = current_time
        i = (current_time - self.start_time - 1) // self.day_to_sec
        if i + 1 > self.n_day:
            return
        self.t[i] += zonal_mean(d.grid_t_c)
        self.teq[i] += zonal_mean(d.grid_t_eq)
        self.u[i] += zonal_mean(d.grid_u_c)
        self.v[i] += zonal_mean(d.grid_v_c)
        self.ps[i] += d.grid_ps_c[0]
        self.count[i] += 1

    def finalize(self):
        c = self.count[:, None, None]
        for f in ("t", "teq", "u", "v", "ps"):
            setattr(self, f, getattr(self, f) / c)
        self.count[:] = 1.0
        s = slice(self.spinup_day, self.n_day)
        self.t_zonal_mean = self.t[s].mean(axis=0)
        self.u_zonal_mean = self.u[s].mean(axis=0)
        self.v_zonal_mean = self.v[s].mean(axis=0)
        self.ps_mean = self.ps[s].mean(axis=0)
