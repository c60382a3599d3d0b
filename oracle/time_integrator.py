"""Filtered leapfrog with implicit hyperdiffusion (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/Atmos_Spectral/Time_Integrator.jl`` (TI):

* ``FilteredLeapfrog``            <- struct + ctor TI:2-38 (damping table TI:31)
* ``update_init_step``            <- ``Update_Init_Step!`` TI:40-43
* ``get_dt`` / ``get_xi``         <- TI:46-56
* ``compute_spectral_damping``    <- ``Compute_Spectral_Damping!`` TI:58-81
* ``filtered_leapfrog``           <- ``Filtered_Leapfrog!`` TI:86-101

Pinned by ``tests/test_oracle_time_integrator.py``: single-mode scalar
recurrences solved in closed form.
"""
import numpy as np


class FilteredLeapfrog:
    def __init__(self, robert_coef, damping_order, damping_coef, eigen,
                 implicit_coef, dt, init_step, start_time, end_time):
        assert damping_order % 2 == 0                      # TI:24
        num_spherical = eigen.shape[1] - 1
        # TI:31: resolution-dependent damping, normalised at degree num_spherical-1
        self.damping = damping_coef * (eigen / eigen[0, num_spherical - 1]) ** damping_order
        self.robert_coef = float(robert_coef)
        self.damping_order = int(damping_order)
        self.damping_coef = float(damping_coef)
        self.laplacian_eigen = eigen
        self.implicit_coef = float(implicit_coef)
        self.dt = int(dt)
        self.init_step = bool(init_step)
        self.time = int(start_time)
        self.start_time, self.end_time = int(start_time), int(end_time)

    def update_init_step(self):
        assert self.init_step                               # TI:41
        self.init_step = False

    def get_dt(self):
        return self.dt if self.init_step else 2 * self.dt

    def get_xi(self):
        return (self.dt * self.implicit_coef if self.init_step
                else 2 * self.dt * self.implicit_coef)

    def compute_spectral_damping(self, Qc, Qp, dQ):
        """TI:76-80; damping table is [m, n], fields are [k, n, m]."""
        d = self.damping.T[None, :, :]
        if self.init_step:
            dQ[...] = (dQ - Qc * d) / (1.0 + self.dt * d)
        else:
            dQ[...] = (dQ - Qp * d) / (1.0 + 2 * self.dt * d)
        return dQ

    def filtered_leapfrog(self, dQ, Qp, Qc, Qn):
        """TI:93-100 (Robert-Asselin filter applied to Qc in place)."""
        r = self.robert_coef
        if self.init_step:
            Qn[...] = Qc + self.dt * dQ
            Qc += r * (-1.0 * Qc + Qn)
        else:
            Qc += r * (Qp - 2 * Qc)
            Qn[...] = Qp + 2 * self.dt * dQ
            Qc += r * Qn
        return Qn
