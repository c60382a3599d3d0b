"""Primitive-equation (Held-Suarez) step and its parts (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, line by line (0-based level indices; half levels 0..nd):

* ``VertCoordinate``          <- Vert_Coordinate.jl:42-56, even sigma :112-117, tables :143-165
* ``vert_advection``          <- ``Vert_Advection!`` Vert_Coordinate.jl:171-210
* ``mass_weighted_global_integral`` <- Vert_Coordinate.jl:215-238
* ``pressure_variables``      <- ``Pressure_Variables!`` Press_And_Geopot.jl:36-67
* ``compute_geopotential``    <- ``Compute_Geopotential!`` Press_And_Geopot.jl:71-111
* ``compute_pressures_and_heights`` <- Press_And_Geopot.jl:3-19
* ``hs_forcing``              <- ``HS_Forcing!`` HS_Forcing.jl:1-130
* ``SemiImplicitSolver``      <- Semi_Implicit.jl:40-188 (matrices, wave matrices)
* ``linear_ps_t_ddiv`` / ``linear_geopot_dt`` / ``linear_geopot_dps`` /
  ``linear_press_gradient_dps`` <- Semi_Implicit.jl:190-341
* ``adjust_dlnps_dt_ddiv``    <- Semi_Implicit.jl:360-404
* ``implicit_correction``     <- Semi_Implicit.jl:462-513
* ``four_in_one``             <- Spectral_Dynamics.jl:88-183
* ``compute_corrections_init`` / ``compute_corrections`` <- Spectral_Dynamics.jl:5-68
* ``spectral_dynamics``       <- Spectral_Dynamics.jl:206-365
* ``spectral_initialize_fields`` <- Spectral_Dynamics.jl:372-449
* ``spectral_dynamics_physics`` / ``atmosphere_update`` <- Spectral_Dynamics.jl:452-490

Pins: tests/test_oracle_primitive.py (isothermal hydrostatic column exact,
Held-Suarez T_eq/k_T closed forms, W_n (I - ξ² eig D) = I, the linear
operators against finite differences of the nonlinear ones they linearise,
an isothermal resting atmosphere that stays at rest without forcing, mass
conservation of the corrections) and tests/test_oracle_hs_pins.py
(``vert_advection``: constant and linear profiles, column budget;
``four_in_one``: pressure-gradient force -R T grad(ln ps) on pure sigma,
omega = 0 for a uniform wind advecting ps, interface mass fluxes;
``compute_corrections_init`` / ``compute_corrections``: mean ps restored and
the new state's mass-weighted energy equal to the target energy computed
independently; each pin is shown to fail on a planted mistake).
"""
import numpy as np

from .dynamics import compute_abs_vor, time_advance


# Compute_Old_Model_Sigma (Vert_Coordinate.jl:157-165) and Compute_V197_Sigma
# (:143-154): b at the half levels, a = 0
SIGMA_TABLES = {
    "mcm": (0.0, 0.03, 0.0707, 0.1311, 0.2102, 0.3036, 0.4062, 0.5138, 0.6226, 0.7284,
            0.8255, 0.9066, 0.9640, 0.9933, 1.0),
    "v197": (0.0, 0.0089163, 0.0342936, 0.0740741, 0.1262002, 0.1886145, 0.2592592,
             0.3360768, 0.4170096, 0.5000000, 0.5829904, 0.6639231, 0.7407407, 0.8113854,
             0.8737997, 0.9259259, 0.9657064, 0.9910837, 1.0),
}


def uneven_sigma(nd, scale_heights, surf_res, exponent, zero_top):
    """Compute_Uneven_Sigma (Vert_Coordinate.jl:120-141) as its docstring
    states it (the code's exp(+...) gives b > 1): ζ = LinRange(1, 0, nd+1),
    b = exp(-(surf_res ζ + (1 - surf_res) ζ^exponent) scale_heights),
    b[nd] = 1, b[0] = 0 when zero_top; a = 0."""
    zeta = np.array([(nd - k) / nd for k in range(nd + 1)])
    z = -(surf_res * zeta + (1.0 - surf_res) * zeta ** exponent)
    b = np.exp(z * scale_heights)
    b[nd] = 1.0
    if zero_top:
        b[0] = 0.0
    return np.zeros(nd + 1), b


def transition(p, p_sigma, p_press):
    """Transition (Vert_Coordinate.jl:93-110) on the σ values p."""
    t = np.zeros(len(p))
    for k, x in enumerate(p):
        if x <= p_press:
            t[k] = 0.0
        elif x >= p_sigma:
            t[k] = 1.0
        else:
            t[k] = np.sin(0.5 * np.pi * (x - p_press) / (p_sigma - p_press)) ** 2
    return t


class VertCoordinate:
    """Vert_Coordinate.jl:42-56 with Compute_Vert_Coord (:61-87): even_sigma
    (:112-117), the fixed tables (their nd must equal their length),
    uneven_sigma (:69-70) and hybrid (:71-77).  The reference cannot run the
    last two (:70 passes a, b before they exist, :72-74 omit nd); they follow
    its docstring and call structure with nd supplied:
    hybrid: (a_σ, b_σ) = uneven σ without a zero top, (b_p, a_p) the same pair
    swapped, trans = Transition(b_σ), a = p_ref (a_σ trans + a_p (1 - trans)),
    b = b_σ trans + b_p (1 - trans)."""

    def __init__(self, nlon, nlat, nd, vert_coord_option, vert_difference_option,
                 vert_advect_scheme, p_ref=101325.0, zero_top=True, scale_heights=4.0,
                 surf_res=1.0, p_press=0.1, p_sigma=0.3, exponent=2.5):
        ak = np.zeros(nd + 1)
        if vert_coord_option == "even_sigma":
            bk = np.linspace(0.0, 1.0, nd + 1)
        elif vert_coord_option in SIGMA_TABLES:
            bk = np.array(SIGMA_TABLES[vert_coord_option])
            if nd != len(bk) - 1:
                raise ValueError(f"{vert_coord_option} has {len(bk) - 1} levels, got nd = {nd}")
        elif vert_coord_option == "uneven_sigma":
            ak, bk = uneven_sigma(nd, scale_heights, surf_res, exponent, zero_top)
        elif vert_coord_option == "hybrid":
            a_sigma, b_sigma = uneven_sigma(nd, scale_heights, surf_res, exponent, False)
            b_press, a_press = uneven_sigma(nd, scale_heights, surf_res, exponent, False)
            trans = transition(b_sigma, p_sigma, p_press)
            ak = p_ref * (a_sigma * trans + a_press * (1.0 - trans))
            bk = b_sigma * trans + b_press * (1.0 - trans)
        else:
            raise ValueError(f"Compute_Vert_Coord {vert_coord_option} is not supported")
        self.nlon, self.nlat, self.nd = nlon, nlat, nd
        self.vert_coord_option = vert_coord_option
        self.vert_difference_option = vert_difference_option
        self.vert_advect_scheme = vert_advect_scheme
        self.p_ref, self.zero_top = p_ref, zero_top
        self.ak = ak
        self.bk = bk
        self.Δak = self.ak[1:] - self.ak[:-1]
        self.Δbk = self.bk[1:] - self.bk[:-1]


def vert_advection(vc, r, dz, w, rdt):
    nd = vc.nd
    flux = np.zeros((nd + 1,) + r.shape[1:])
    if vc.vert_advect_scheme == "second_centered_wts":
        flux[1:nd] = w[1:nd] * (r[0:nd - 1] + (r[1:nd] - r[0:nd - 1]) * dz[0:nd - 1]
                                / (dz[0:nd - 1] + dz[1:nd]))
    elif vc.vert_advect_scheme == "second_centered":
        flux[1:nd] = w[1:nd] * (r[0:nd - 1] + r[1:nd]) / 2.0
    else:
        raise ValueError(vc.vert_advect_scheme)
    rdt[...] = (flux[0:nd] - flux[1:nd + 1] + r * (w[1:nd + 1] - w[0:nd])) / dz
    return rdt


def mass_weighted_global_integral(vc, mesh, atmo, grid_data, grid_ps):
    vi = np.zeros_like(grid_ps)
    for k in range(vc.nd):
        dp = vc.Δak[k] + vc.Δbk[k] * grid_ps
        vi += grid_data[k:k + 1] * dp
    return mesh.area_weighted_global_mean(vi) / atmo.grav


def pressure_variables(vc, grid_ps, p_half, Δp, lnp_half, p_full, lnp_full):
    nd = vc.nd
    for k in range(nd + 1):
        p_half[k] = vc.ak[k] + vc.bk[k] * grid_ps[0]
    Δp[...] = p_half[1:nd + 1] - p_half[0:nd]
    kt = 1 if vc.zero_top else 0
    lnp_half[kt:] = np.log(p_half[kt:])
    lnp_full[kt:nd] = (lnp_half[kt + 1:nd + 1] + p_half[kt:nd] * (lnp_half[kt + 1:nd + 1] - lnp_half[kt:nd])
                       / Δp[kt:nd] - 1.0)
    if vc.zero_top:
        lnp_half[0] = 0.0
        lnp_full[0] = lnp_half[1] - 1.0
    p_full[...] = np.exp(lnp_full)


def compute_geopotential(vc, atmo, lnp_half, lnp_full, grid_t, geopots, geopot_full, geopot_half):
    nd = vc.nd
    geopot_half[nd] = geopots[0]
    if vc.zero_top:
        kt = 1
        geopot_half[0] = 0.0
    else:
        kt = 0
    for k in range(nd - 1, kt - 1, -1):
        geopot_half[k] = geopot_half[k + 1] + atmo.rdgas * grid_t[k] * (lnp_half[k + 1] - lnp_half[k])
    for k in range(nd):
        geopot_full[k] = geopot_half[k + 1] + atmo.rdgas * grid_t[k] * (lnp_half[k + 1] - lnp_full[k])


def compute_pressures_and_heights(atmo, vc, grid_ps, geopots, grid_t, p_half, Δp, lnp_half,
                                  p_full, lnp_full, z_full, z_half):
    pressure_variables(vc, grid_ps, p_half, Δp, lnp_half, p_full, lnp_full)
    compute_geopotential(vc, atmo, lnp_half, lnp_full, grid_t, geopots, z_full, z_half)
    z_full /= atmo.grav
    z_half /= atmo.grav


def hs_forcing(atmo, dt, sin_theta, u, v, p_half, p_full, t, δu, δv, t_eq, δt, params):
    """HS_Forcing.jl:1-33 with Rayleigh_Damping! (:101-130) and
    Newtonian_Damping! (:37-97)."""
    nd = u.shape[0]
    σ_b, k_f, k_a, k_s = params["σ_b"], params["k_f"], params["k_a"], params["k_s"]
    ΔT_y, Δθ_z = params["ΔT_y"], params["Δθ_z"]
    day = 86400
    kf = k_f / day
    ps = p_half[nd]
    for k in range(nd):
        σ = p_full[k] / ps
        assert np.max(σ - σ_b) / (1.0 - σ_b) <= 1.0
        k_v = kf / (1.0 - σ_b) * np.maximum(0, σ - σ_b)
        δu[k] = -k_v * u[k]
        δv[k] = -k_v * v[k]
    δt[...] = -((u + 0.5 * δu * dt) * δu + (v + 0.5 * δv * dt) * δv) / atmo.cp_air
    t_zero, t_strat = 315.0, 200.0
    ka, ks = k_a / day, k_s / day
    p_ref = 1.0e5
    s2 = (sin_theta * sin_theta)[:, None]
    c2 = 1.0 - s2
    c4 = c2 * c2
    for k in range(nd):
        pn = p_full[k] / p_ref
        te = (t_zero - ΔT_y * s2 - Δθ_z * c2 * np.log(pn)) * pn ** atmo.kappa
        t_eq[k] = np.maximum(t_strat, te)
    for k in range(nd):
        σ = p_full[k] / ps
        assert np.max(σ - σ_b) / (1.0 - σ_b) <= 1.0
        k_t = ka + (ks - ka) / (1.0 - σ_b) * np.maximum(0.0, σ - σ_b) * c4
        δt[k] -= k_t * (t[k] - t_eq[k])


# ------------------------------------------------------------ semi-implicit
def linear_ps_t_ddiv(vc, atmo, ddiv, ps_ref, Δp_ref, lnp_half_ref, lnp_full_ref, t_ref,
                     M_half, Mdt_half, dps, dt):
    """Semi_Implicit.jl:271-341.  ddiv [nd, ...] -> dps [1, ...], dt [nd, ...]."""
    nd = vc.nd
    kappa = atmo.kappa
    dmean_tot = np.zeros(ddiv.shape[1:], ddiv.dtype)
    for k in range(nd):
        assert np.isclose(vc.Δak[k] + vc.Δbk[k] * ps_ref, Δp_ref[k])
        Δlnp_p = lnp_half_ref[k + 1] - lnp_full_ref[k]
        Δlnp = lnp_half_ref[k + 1] - lnp_half_ref[k]
        dmean = ddiv[k] * Δp_ref[k]
        dt[k] = -kappa * t_ref[k] * (dmean_tot * Δlnp + dmean * Δlnp_p) / Δp_ref[k]
        dmean_tot = dmean_tot + dmean
        M_half[k + 1] = -dmean_tot
    dps[0] = -dmean_tot
    for k in range(nd - 1):
        M_half[k + 1] += dmean_tot * vc.bk[k + 1]
    M_half[0] = 0.0
    M_half[nd] = 0.0
    for k in range(1, nd):
        Mdt_half[k] = M_half[k] * (t_ref[k] - t_ref[k - 1])
    Mdt_half[0] = 0.0
    Mdt_half[nd] = 0.0
    for k in range(nd):
        dt[k] -= 0.5 * (Mdt_half[k + 1] + Mdt_half[k]) / Δp_ref[k]


def linear_geopot_dt(vc, atmo, lnp_half_ref, lnp_full_ref, dt, t_ref, geopot_half, geopot):
    """Semi_Implicit.jl:213-241."""
    nd = vc.nd
    R = atmo.rdgas
    geopot_half[nd] = 0.0
    for k in range(nd - 1, 0, -1):
        geopot_half[k] = geopot_half[k + 1] + R * (dt[k] * (lnp_half_ref[k + 1] - lnp_half_ref[k]))
    for k in range(nd):
        geopot[k] = geopot_half[k + 1] + R * (dt[k] * (lnp_half_ref[k + 1] - lnp_full_ref[k]))


def linear_geopot_dps(vc, atmo, dlnp_half_ref, dlnp_full_ref, lnp_half_ref, lnp_full_ref, t_ref):
    """Semi_Implicit.jl:243-268."""
    nd = vc.nd
    R = atmo.rdgas
    gh = np.zeros(nd + 1)
    g = np.zeros(nd)
    for k in range(nd - 1, 0, -1):
        gh[k] = gh[k + 1] + R * (t_ref[k] * (dlnp_half_ref[k + 1] - dlnp_half_ref[k]))
    for k in range(nd):
        g[k] = gh[k + 1] + R * (t_ref[k] * (dlnp_half_ref[k + 1] - dlnp_full_ref[k]))
    return g


def linear_press_gradient_dps(vc, atmo, ps_ref, Δp_ref, lnp_half_ref, lnp_full_ref, t_ref):
    """Semi_Implicit.jl:190-210."""
    nd, bk = vc.nd, vc.bk
    out = np.zeros(nd)
    for k in range(nd):
        Δlnp_p = lnp_half_ref[k + 1] - lnp_full_ref[k]
        Δlnp_m = lnp_full_ref[k] - lnp_half_ref[k]
        out[k] = atmo.rdgas * t_ref[k] * (bk[k + 1] * Δlnp_p + bk[k] * Δlnp_m) / Δp_ref[k]
    return out


def build_wave_matrices(num_wavenumbers, eigen, div_mat, xi):
    """Semi_Implicit.jl:165-180."""
    nd = div_mat.shape[0]
    W = np.zeros((num_wavenumbers + 1, nd, nd))
    for i in range(num_wavenumbers + 1):
        factor = xi ** 2 * eigen[0, i]
        W[i] = np.linalg.inv(np.eye(nd) - factor * div_mat)
    return W


class SemiImplicitSolver:
    """Semi_Implicit.jl:40-97 and Build_Implicit_Matrices (:112-160)."""

    def __init__(self, vc, atmo, integrator, ps_ref, t_ref, wave_numbers):
        nd = vc.nd
        self.integrator, self.nd = integrator, nd
        self.ps_ref = ps_ref
        self.t_ref = np.asarray(t_ref, float)
        ps3 = np.full((1, 1, 1), ps_ref)
        ph, Δp, lh, pf, lf = (np.zeros((nd + 1, 1, 1)), np.zeros((nd, 1, 1)), np.zeros((nd + 1, 1, 1)),
                              np.zeros((nd, 1, 1)), np.zeros((nd, 1, 1)))
        pressure_variables(vc, ps3, ph, Δp, lh, pf, lf)
        self.Δp_ref, self.p_half_ref = Δp[:, 0, 0], ph[:, 0, 0]
        self.lnp_half_ref, self.lnp_full_ref = lh[:, 0, 0], lf[:, 0, 0]
        with np.errstate(invalid="ignore", divide="ignore"):
            dlh = vc.bk / self.p_half_ref          # index 0 is 0/0 and is never used
        dlf = np.zeros(nd)
        for k in range(nd):
            Δlnp_p = self.lnp_half_ref[k + 1] - self.lnp_full_ref[k]
            Δlnp_m = self.lnp_full_ref[k] - self.lnp_half_ref[k]
            dlf[k] = (vc.bk[k + 1] * Δlnp_p + vc.bk[k] * Δlnp_m) / self.Δp_ref[k]
        tau, gamma, nu = np.zeros((nd, nd)), np.zeros((nd, nd)), np.zeros(nd)
        Mh, Mdt = np.zeros((nd + 1, 1, 1), complex), np.zeros((nd + 1, 1, 1), complex)
        dps, dt = np.zeros((1, 1, 1), complex), np.zeros((nd, 1, 1), complex)
        gh, g = np.zeros((nd + 1, 1, 1), complex), np.zeros((nd, 1, 1), complex)
        for k in range(nd):
            inp = np.zeros((nd, 1, 1), complex)
            inp[k] = 1.0
            linear_ps_t_ddiv(vc, atmo, inp, ps_ref, self.Δp_ref, self.lnp_half_ref,
                             self.lnp_full_ref, self.t_ref, Mh, Mdt, dps, dt)
            nu[k] = -dps[0, 0, 0].real
            tau[:, k] = -dt[:, 0, 0].real
            linear_geopot_dt(vc, atmo, self.lnp_half_ref, self.lnp_full_ref, inp, self.t_ref, gh, g)
            gamma[:, k] = g[:, 0, 0].real
        h1 = linear_press_gradient_dps(vc, atmo, ps_ref, self.Δp_ref, self.lnp_half_ref,
                                       self.lnp_full_ref, self.t_ref)
        h2 = linear_geopot_dps(vc, atmo, dlh, dlf, self.lnp_half_ref, self.lnp_full_ref, self.t_ref)
        self.h = h1 + h2
        self.div_mat = np.outer(self.h, nu) + gamma @ tau
        self.tau, self.gamma, self.nu = tau, gamma, nu
        self.wave_numbers = wave_numbers
        self.num_wavenumbers = wave_numbers.shape[1] - 2
        self.wave_matrix = build_wave_matrices(self.num_wavenumbers, integrator.laplacian_eigen,
                                               self.div_mat, integrator.get_xi())

    def update_init_step(self):
        """Semi_Implicit.jl:182-188."""
        self.integrator.update_init_step()
        self.wave_matrix = build_wave_matrices(self.num_wavenumbers, self.integrator.laplacian_eigen,
                                               self.div_mat, self.integrator.get_xi())


def adjust_dlnps_dt_ddiv(si, vc, atmo, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt):
    """Semi_Implicit.jl:360-404."""
    nd = vc.nd
    shp = ddiv.shape[1:]
    Mh, Mdt = np.zeros((nd + 1,) + shp, complex), np.zeros((nd + 1,) + shp, complex)
    dps_tmp, dt_tmp = np.zeros((1,) + shp, complex), np.zeros((nd,) + shp, complex)
    ddiv_tmp = div_p - div_c
    linear_ps_t_ddiv(vc, atmo, ddiv_tmp, si.ps_ref, si.Δp_ref, si.lnp_half_ref, si.lnp_full_ref,
                     si.t_ref, Mh, Mdt, dps_tmp, dt_tmp)
    dt += dt_tmp
    dlnps += dps_tmp / si.ps_ref
    xi = si.integrator.get_xi()
    dt_tmp = t_p - t_c + xi * dt
    dlnps_tmp = lnps_p - lnps_c + xi * dlnps
    gh, g = np.zeros((nd + 1,) + shp, complex), np.zeros((nd,) + shp, complex)
    linear_geopot_dt(vc, atmo, si.lnp_half_ref, si.lnp_full_ref, dt_tmp, si.t_ref, gh, g)
    eig = si.integrator.laplacian_eigen.T
    for k in range(nd):
        ddiv[k] -= eig * (g[k] + si.h[k] * si.ps_ref * dlnps_tmp[0])


def implicit_correction(si, vc, atmo, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt):
    """Semi_Implicit.jl:462-513."""
    xi = si.integrator.get_xi()
    adjust_dlnps_dt_ddiv(si, vc, atmo, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt)
    nf = ddiv.shape[2]
    ns = ddiv.shape[1]
    for m in range(nf):
        for n in range(m, ns):
            if n <= si.num_wavenumbers:
                ddiv[:, n, m] = si.wave_matrix[n] @ ddiv[:, n, m]
    nd = vc.nd
    shp = ddiv.shape[1:]
    Mh, Mdt = np.zeros((nd + 1,) + shp, complex), np.zeros((nd + 1,) + shp, complex)
    dps_tmp, dt_tmp = np.zeros((1,) + shp, complex), np.zeros((nd,) + shp, complex)
    linear_ps_t_ddiv(vc, atmo, ddiv, si.ps_ref, si.Δp_ref, si.lnp_half_ref, si.lnp_full_ref,
                     si.t_ref, Mh, Mdt, dps_tmp, dt_tmp)
    dt += xi * dt_tmp
    dlnps += xi / si.ps_ref * dps_tmp


# ------------------------------------------------------------ grid columns
def four_in_one(vc, atmo, div, u, v, ps, Δp, lnp_half, lnp_full, p_full, dλ_ps, dθ_ps, t,
                M_half, w_full, δu, δv, δps, δt):
    """Spectral_Dynamics.jl:88-183."""
    R = atmo.rdgas
    nd, bk, Δbk = vc.nd, vc.bk, vc.Δbk
    kappa = R / atmo.cp_air
    dmean_tot = np.zeros(ps.shape[1:])
    for k in range(nd):
        dp = Δp[k]
        Δlnp_p = lnp_half[k + 1] - lnp_full[k]
        Δlnp_m = lnp_full[k] - lnp_half[k]
        Δlnp = lnp_half[k + 1] - lnp_half[k]
        x1 = (bk[k] * Δlnp_m + bk[k + 1] * Δlnp_p) / dp
        dl = x1 * dλ_ps[0]
        dth = x1 * dθ_ps[0]
        δu[k] -= R * t[k] * dl
        δv[k] -= R * t[k] * dth
        dmean = div[k] * dp + Δbk[k] * (u[k] * dλ_ps[0] + v[k] * dθ_ps[0])
        x5 = -(dmean_tot * Δlnp + dmean * Δlnp_p) / dp + u[k] * dl + v[k] * dth
        δt[k] += kappa * t[k] * x5
        w_full[k] = x5 * p_full[k]
        dmean_tot = dmean_tot + dmean
        M_half[k + 1] = -dmean_tot
    δps[0] -= dmean_tot
    for k in range(nd - 1):
        M_half[k + 1] += dmean_tot * bk[k + 1]
    M_half[0] = 0.0
    M_half[nd] = 0.0


def compute_corrections_init(vc, mesh, atmo, u_p, v_p, ps_p, t_p, δu, δv, δt, dt, energy_tmp):
    """Spectral_Dynamics.jl:5-28."""
    mean_ps_p = mean_energy_p = None
    if atmo.do_mass_correction:
        mean_ps_p = mesh.area_weighted_global_mean(ps_p)
    if atmo.do_energy_correction:
        energy_tmp[...] = (0.5 * ((u_p + dt * δu) ** 2 + (v_p + dt * δv) ** 2)
                           + atmo.cp_air * (t_p + dt * δt))
        mean_energy_p = mass_weighted_global_integral(vc, mesh, atmo, energy_tmp, ps_p)
    if atmo.do_water_correction:
        raise RuntimeError("water correction has not implemented")
    return mean_ps_p, mean_energy_p


def compute_corrections(vc, mesh, atmo, mean_ps_p, mean_energy_p, u_n, v_n, energy_tmp, ps_n,
                        spe_lnps_n, t_n, spe_t_n):
    """Spectral_Dynamics.jl:30-68."""
    if atmo.do_mass_correction:
        mean_ps_n = mesh.area_weighted_global_mean(ps_n)
        factor = mean_ps_p / mean_ps_n
        ps_n *= factor
        spe_lnps_n[0, 0, 0] += np.log(factor)
    if atmo.do_energy_correction:
        energy_tmp[...] = 0.5 * (u_n ** 2 + v_n ** 2) + atmo.cp_air * t_n
        mean_energy_n = mass_weighted_global_integral(vc, mesh, atmo, energy_tmp, ps_n)
        corr = atmo.grav * (mean_energy_p - mean_energy_n) / (atmo.cp_air * mean_ps_p)
        t_n += corr
        spe_t_n[:, 0, 0] += corr
    if atmo.do_water_correction:
        raise RuntimeError("water correction has not implemented")


def spectral_dynamics(mesh, vc, atmo, d, si):
    """Spectral_Dynamics.jl:206-365."""
    integ = si.integrator
    dt = integ.get_dt()
    mean_ps_p, mean_energy_p = compute_corrections_init(
        vc, mesh, atmo, d.grid_u_p, d.grid_v_p, d.grid_ps_p, d.grid_t_p, d.grid_δu, d.grid_δv,
        d.grid_δt, dt, d.grid_energy_full)
    pressure_variables(vc, d.grid_ps_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half, d.grid_p_full,
                       d.grid_lnp_full)
    mesh.compute_gradients(d.spe_lnps_c, d.grid_dλ_ps, d.grid_dθ_ps)
    d.grid_dλ_ps *= d.grid_ps_c
    d.grid_dθ_ps *= d.grid_ps_c
    four_in_one(vc, atmo, d.grid_div, d.grid_u_c, d.grid_v_c, d.grid_ps_c, d.grid_Δp,
                d.grid_lnp_half, d.grid_lnp_full, d.grid_p_full, d.grid_dλ_ps, d.grid_dθ_ps,
                d.grid_t_c, d.grid_M_half, d.grid_w_full, d.grid_δu, d.grid_δv, d.grid_δps,
                d.grid_δt)
    compute_geopotential(vc, atmo, d.grid_lnp_half, d.grid_lnp_full, d.grid_t_c, d.grid_geopots,
                         d.grid_geopot_full, d.grid_geopot_half)
    d.grid_δlnps[...] = d.grid_δps / d.grid_ps_c
    mesh.trans_grid_to_spherical(d.grid_δlnps, d.spe_δlnps)
    dQ = d.grid_d_full1
    vert_advection(vc, d.grid_u_c, d.grid_Δp, d.grid_M_half, dQ)
    d.grid_δu += dQ
    vert_advection(vc, d.grid_v_c, d.grid_Δp, d.grid_M_half, dQ)
    d.grid_δv += dQ
    vert_advection(vc, d.grid_t_c, d.grid_Δp, d.grid_M_half, dQ)
    d.grid_δt += dQ
    mesh.add_horizontal_advection(d.spe_t_c, d.grid_u_c, d.grid_v_c, d.grid_δt)
    mesh.trans_grid_to_spherical(d.grid_δt, d.spe_δt)
    compute_abs_vor(d.grid_vor, atmo.coriolis, d.grid_absvor)
    d.grid_δu += d.grid_absvor * d.grid_v_c
    d.grid_δv -= d.grid_absvor * d.grid_u_c
    mesh.vor_div_from_grid_uv(d.grid_δu, d.grid_δv, d.spe_δvor, d.spe_δdiv)
    d.grid_energy_full[...] = d.grid_geopot_full + 0.5 * (d.grid_u_c ** 2 + d.grid_v_c ** 2)
    mesh.trans_grid_to_spherical(d.grid_energy_full, d.spe_energy)
    mesh.apply_laplacian(d.spe_energy)
    d.spe_δdiv -= d.spe_energy
    implicit_correction(si, vc, atmo, d.spe_div_c, d.spe_div_p, d.spe_lnps_c, d.spe_lnps_p,
                        d.spe_t_c, d.spe_t_p, d.spe_δdiv, d.spe_δlnps, d.spe_δt)
    integ.compute_spectral_damping(d.spe_vor_c, d.spe_vor_p, d.spe_δvor)
    integ.compute_spectral_damping(d.spe_div_c, d.spe_div_p, d.spe_δdiv)
    integ.compute_spectral_damping(d.spe_t_c, d.spe_t_p, d.spe_δt)
    integ.filtered_leapfrog(d.spe_δvor, d.spe_vor_p, d.spe_vor_c, d.spe_vor_n)
    integ.filtered_leapfrog(d.spe_δdiv, d.spe_div_p, d.spe_div_c, d.spe_div_n)
    integ.filtered_leapfrog(d.spe_δlnps, d.spe_lnps_p, d.spe_lnps_c, d.spe_lnps_n)
    integ.filtered_leapfrog(d.spe_δt, d.spe_t_p, d.spe_t_c, d.spe_t_n)
    mesh.trans_spherical_to_grid(d.spe_vor_n, d.grid_vor)
    mesh.trans_spherical_to_grid(d.spe_div_n, d.grid_div)
    mesh.uv_grid_from_vor_div(d.spe_vor_n, d.spe_div_n, d.grid_u_n, d.grid_v_n)
    mesh.trans_spherical_to_grid(d.spe_t_n, d.grid_t_n)
    mesh.trans_spherical_to_grid(d.spe_lnps_n, d.grid_lnps)
    d.grid_ps_n[...] = np.exp(d.grid_lnps)
    compute_corrections(vc, mesh, atmo, mean_ps_p, mean_energy_p, d.grid_u_n, d.grid_v_n,
                        d.grid_energy_full, d.grid_ps_n, d.spe_lnps_n, d.grid_t_n, d.spe_t_n)
    time_advance(d)
    pressure_variables(vc, d.grid_ps_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half, d.grid_p_full,
                       d.grid_lnp_full)


def spectral_initialize_fields(mesh, atmo, vc, sea_level_ps_ref, init_t, geopots, d,
                               grid_t_noise=None):
    """Spectral_Dynamics.jl:372-449.  ``grid_t_noise`` (optional) is added to
    the isothermal grid temperature before the spectral filtering, the
    seeded-noise initial state of the Held-Suarez benchmark configuration."""
    d.grid_t_c[...] = init_t
    if grid_t_noise is not None:
        d.grid_t_c += grid_t_noise
    d.grid_lnps[...] = np.log(sea_level_ps_ref) - geopots / (atmo.rdgas * init_t)
    d.grid_ps_c[...] = np.exp(d.grid_lnps)
    d.spe_div_c[...] = 0.0
    d.spe_vor_c[...] = 0.0
    nd = mesh.nd
    pert = 1.0e-7 / np.sqrt(2.0)
    for k in range(nd - 3, nd):
        d.spe_vor_c[k, 4, 1] = pert
        d.spe_vor_c[k, 8, 5] = pert
        d.spe_vor_c[k, 3, 1] = pert
        d.spe_vor_c[k, 7, 5] = pert
    mesh.uv_grid_from_vor_div(d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    mesh.trans_grid_to_spherical(d.grid_t_c, d.spe_t_c)
    mesh.trans_spherical_to_grid(d.spe_t_c, d.grid_t_c)
    mesh.trans_grid_to_spherical(d.grid_lnps, d.spe_lnps_c)
    mesh.trans_spherical_to_grid(d.spe_lnps_c, d.grid_lnps)
    d.grid_ps_c[...] = np.exp(d.grid_lnps)
    mesh.vor_div_from_grid_uv(d.grid_u_c, d.grid_v_c, d.spe_vor_c, d.spe_div_c)
    mesh.uv_grid_from_vor_div(d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    mesh.trans_spherical_to_grid(d.spe_vor_c, d.grid_vor)
    mesh.trans_spherical_to_grid(d.spe_div_c, d.grid_div)
    pressure_variables(vc, d.grid_ps_c, d.grid_p_half, d.grid_Δp, d.grid_lnp_half, d.grid_p_full,
                       d.grid_lnp_full)
    for f in ("vor", "div", "lnps", "t"):
        getattr(d, f"spe_{f}_p")[...] = getattr(d, f"spe_{f}_c")
    for f in ("u", "v", "ps", "t"):
        getattr(d, f"grid_{f}_p")[...] = getattr(d, f"grid_{f}_c")


def spectral_dynamics_physics(atmo, mesh, d, dt, params):
    """Spectral_Dynamics.jl:452-467."""
    d.grid_δps[...] = 0.0
    hs_forcing(atmo, dt, mesh.sin_theta, d.grid_u_p, d.grid_v_p, d.grid_p_half, d.grid_p_full,
               d.grid_t_p, d.grid_δu, d.grid_δv, d.grid_t_eq, d.grid_δt, params)


def atmosphere_update(mesh, atmo, vc, si, d, params):
    """Spectral_Dynamics.jl:470-490."""
    dt = si.integrator.get_dt()
    spectral_dynamics_physics(atmo, mesh, d, dt, params)
    spectral_dynamics(mesh, vc, atmo, d, si)
    compute_pressures_and_heights(atmo, vc, d.grid_ps_c, d.grid_geopots, d.grid_t_c, d.grid_p_half,
                                  d.grid_Δp, d.grid_lnp_half, d.grid_p_full, d.grid_lnp_full,
                                  d.grid_z_full, d.grid_z_half)
