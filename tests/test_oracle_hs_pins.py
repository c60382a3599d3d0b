"""Independent pins for three parts of the Held-Suarez oracle step (CPU).

Reference: ``Vert_Advection!`` Vert_Coordinate.jl:171-210, ``Four_In_One!``
Spectral_Dynamics.jl:88-183 (the pressure-gradient weights x1, :127-136, and
the mass-flux / omega terms, :139-163) and the energy and mass corrections,
``Compute_Corrections_Init`` / ``Compute_Corrections!`` Spectral_Dynamics.jl:5-68.

Every expected value here comes from the mathematics the discretisation is
built to satisfy, not from a second copy of the oracle's formulas:

* vertical advection: a constant profile is not changed; a profile linear in
  the height coordinate gets rdt_k = -1/2 (w_{k-1/2} + w_{k+1/2}) dr/dz for
  the weighted scheme on any spacing (and the plain scheme on uniform
  spacing); the column budget sum_k rdt_k dz_k = sum_k r_k (w_{k+1/2} -
  w_{k-1/2}) holds because the boundary fluxes vanish;
* Four_In_One!: on a pure sigma coordinate the Simmons-Burridge weights give
  grad(p)/p = grad(ps)/ps at every full level, so the pressure-gradient force
  is -R T grad(ln ps); a horizontally uniform, divergence-free wind that
  only advects ps leaves dp/dt = 0 along a trajectory (omega = 0, M = 0 at
  every interface, no heating) and gives dps/dt = -v.grad(ps); with no
  pressure gradient the interface mass flux is the partial column
  divergence corrected by b times the total;
* corrections: after Compute_Corrections!, the global mean of ps equals the
  previous one, and the mass-weighted energy 1/2|v|^2 + cp T of the new state
  equals the mass-weighted energy of (u_p + dt du, v_p + dt dv, T_p + dt dT)
  over ps_p, computed here with numpy's Gauss-Legendre weights; the (0,0)
  spectral coefficients move by what the grid constants moved.

The last test plants the four mistakes a review found the older pins blind
to (a flipped r dw sign, swapped b weights in x1, the energy correction
without g, and the energy target built from u_p instead of u_p + dt du) into
copies of the oracle functions and checks that each fails one of these pins.
"""
import inspect
import types

import numpy as np
import pytest

from oracle import SpectralSphericalMesh
from oracle import dynamics as odyn
from oracle import primitive as P

R = 287.04
COORDS = [("even_sigma", 10), ("mcm", 14), ("v197", 18), ("uneven_sigma", 20)]


# ------------------------------------------------------------ vertical advection
def _vc(nd, scheme, coord="even_sigma"):
    return P.VertCoordinate(1, 1, nd, coord, "simmons_and_burridge", scheme, 1e5)


def _w(rng, nd, shape):
    w = rng.standard_normal((nd + 1,) + shape)
    w[0] = 0.0          # no flux through the top and the bottom (M_half's boundary values)
    w[nd] = 0.0
    return w


def pin_vert_advection(vert_advection=P.vert_advection):
    """Largest relative violation of the three vertical-advection pins."""
    rng = np.random.default_rng(11)
    nd, shape = 12, (3, 5)
    errs = []
    for scheme, uniform in (("second_centered_wts", False), ("second_centered_wts", True),
                            ("second_centered", True)):
        vc = _vc(nd, scheme)
        if uniform:
            dz = np.full((nd,) + shape, 0.7)
        else:
            dz = 0.2 + rng.random((nd,) + shape)
        w = _w(rng, nd, shape)
        # constant profile: no tendency
        rdt = np.zeros((nd,) + shape)
        vert_advection(vc, np.full((nd,) + shape, 3.25), dz, w, rdt)
        errs.append(np.abs(rdt).max() / 3.25)
        # linear profile r = a + b z at the cell centres
        zc = np.cumsum(dz, axis=0) - 0.5 * dz
        a, b = 1.5, -0.8
        vert_advection(vc, a + b * zc, dz, w, rdt)
        expect = -0.5 * (w[:nd] + w[1:]) * b
        errs.append(np.abs(rdt - expect).max() / np.abs(expect).max())
        # column budget for an arbitrary profile
        r = rng.standard_normal((nd,) + shape)
        vert_advection(vc, r, dz, w, rdt)
        lhs = (rdt * dz).sum(axis=0)
        rhs = (r * (w[1:] - w[:nd])).sum(axis=0)
        errs.append(np.abs(lhs - rhs).max() / np.abs(rhs).max())
    return max(errs)


def test_vert_advection_pins():
    assert pin_vert_advection() < 1e-13


# ------------------------------------------------------------ Four_In_One!
def _columns(coord, nd, ps):
    """Pressure variables of pure-sigma columns over a ps field [1, ny, nx]."""
    vc = P.VertCoordinate(ps.shape[2], ps.shape[1], nd, coord, "simmons_and_burridge",
                          "second_centered_wts", 1e5)
    sh = lambda k: np.zeros((k,) + ps.shape[1:])
    ph, Δp, lh, pf, lf = sh(nd + 1), sh(nd), sh(nd + 1), sh(nd), sh(nd)
    P.pressure_variables(vc, ps, ph, Δp, lh, pf, lf)
    return vc, ph, Δp, lh, pf, lf


def _atmo(nd):
    return odyn.AtmoData("hs", 1, 1, nd, False, False, False, False, np.zeros(1), 6.371e6)


def pin_four_in_one(four_in_one=P.four_in_one):
    """Largest relative violation of the Four_In_One! pins over the three
    pure sigma coordinates."""
    rng = np.random.default_rng(5)
    errs = []
    for coord, nd in COORDS:
        shape = (4, 6)
        ps = 9.0e4 + 2.0e4 * rng.random((1,) + shape)
        vc, ph, Δp, lh, pf, lf = _columns(coord, nd, ps)
        atmo = _atmo(nd)
        kappa = R / atmo.cp_air
        T = 250.0 + 30.0 * rng.random((nd,) + shape)
        dλps = 1.0e-2 * rng.standard_normal((1,) + shape)     # grad(ps) components
        dθps = 1.0e-2 * rng.standard_normal((1,) + shape)
        z = lambda k: np.zeros((k,) + shape)

        # 1. pressure-gradient force -R T grad(ln ps), whatever the winds and divergence
        u, v, div = (rng.standard_normal((nd,) + shape) for _ in range(3))
        δu0, δv0 = rng.standard_normal((2, nd) + shape)
        δu, δv = δu0.copy(), δv0.copy()
        four_in_one(vc, atmo, 1e-5 * div, 10 * u, 10 * v, ps, Δp, lh, lf, pf, dλps, dθps, T,
                    z(nd + 1), z(nd), δu, δv, z(1), z(nd))
        for got, base, g in ((δu, δu0, dλps), (δv, δv0, dθps)):
            expect = -R * T * g / ps
            errs.append(np.abs((got - base) - expect).max() / np.abs(expect).max())

        # 2. uniform divergence-free wind advecting ps: omega = 0, M = 0, no heating
        u = np.broadcast_to(rng.standard_normal((1,) + shape) * 20, (nd,) + shape).copy()
        v = np.broadcast_to(rng.standard_normal((1,) + shape) * 20, (nd,) + shape).copy()
        Mh, w_full, δps, δt = z(nd + 1), z(nd), z(1), z(nd)
        four_in_one(vc, atmo, z(nd), u, v, ps, Δp, lh, lf, pf, dλps, dθps, T, Mh, w_full,
                    z(nd), z(nd), δps, δt)
        adv = u[0] * dλps[0] + v[0] * dθps[0]
        scale = np.abs(adv).max()
        errs.append(np.abs(δps[0] + adv).max() / scale)
        errs.append(np.abs(Mh).max() / scale)
        errs.append(np.abs(w_full / pf).max() / (scale / ps.min()))
        errs.append(np.abs(δt).max() / (kappa * T.max() * scale / ps.min()))

        # 3. divergence with no pressure gradient: column mass budget
        div = 1e-5 * rng.standard_normal((nd,) + shape)
        Mh, δps = z(nd + 1), z(1)
        four_in_one(vc, atmo, div, z(nd), z(nd), ps, Δp, lh, lf, pf, z(1), z(1), T, Mh, z(nd),
                    z(nd), z(nd), δps, z(nd))
        flux = div * Δp                          # mass divergence of each layer
        tot = flux.sum(axis=0)
        errs.append(np.abs(δps[0] + tot).max() / np.abs(tot).max())
        part = np.cumsum(flux, axis=0)           # layers above each interface
        expect = -part[:-1] + vc.bk[1:nd, None, None] * tot
        errs.append(np.abs(Mh[1:nd] - expect).max() / np.abs(part).max())
        errs.append(max(np.abs(Mh[0]).max(), np.abs(Mh[nd]).max()))
    return max(errs)


def test_four_in_one_pins():
    assert pin_four_in_one() < 1e-12


# ------------------------------------------------------------ corrections
def _gauss_weights(nlat):
    """Gauss-Legendre weights (numpy), independent of the oracle's Newton iteration."""
    return np.polynomial.legendre.leggauss(nlat)[1]


def _global_mean(f2d, wts):
    """Area-weighted global mean over a [nlat, nlon] field (weights sum to 2)."""
    return float((f2d.mean(axis=1) * wts).sum() / 2.0)


def pin_corrections(init=P.compute_corrections_init, correct=P.compute_corrections):
    """Largest relative violation of the mass and energy correction pins."""
    rng = np.random.default_rng(7)
    N, nlat = 21, 32
    nlon = 2 * nlat
    errs = []
    for coord, nd in COORDS:
        mesh = SpectralSphericalMesh(N, N + 1, nlat, nlon, nd, 6.371e6)
        vc = P.VertCoordinate(nlon, nlat, nd, coord, "simmons_and_burridge", "second_centered_wts", 1e5)
        atmo = odyn.AtmoData("hs", nlon, nlat, nd, True, True, False, False, mesh.sin_theta, 6.371e6)
        g3 = lambda k, s=1.0: s * rng.standard_normal((k, nlat, nlon))
        u_p, v_p = g3(nd, 15.0), g3(nd, 15.0)
        t_p = 260.0 + g3(nd, 10.0)
        ps_p = 1.0e5 + g3(1, 500.0)
        # tendencies shaped like the forcing's: Rayleigh friction and cooling
        kf = 1.0 / 86400.0
        δu, δv = -kf * u_p + g3(nd, 1e-5), -kf * v_p + g3(nd, 1e-5)
        δt = -(t_p - 250.0) / (40 * 86400.0)
        dt = 1200
        u_n, v_n = u_p + g3(nd, 0.5), v_p + g3(nd, 0.5)
        t_n0 = t_p + g3(nd, 0.3)
        ps_n0 = ps_p * (1.0 + 1e-3) + g3(1, 50.0)
        t_n, ps_n = t_n0.copy(), ps_n0.copy()
        spe_lnps_n = rng.standard_normal((1, N + 2, N + 1)) + 0j
        spe_t_n = rng.standard_normal((nd, N + 2, N + 1)) + 0j
        spe_lnps_0, spe_t_0 = spe_lnps_n.copy(), spe_t_n.copy()
        tmp = np.zeros((nd, nlat, nlon))
        mps, me = init(vc, mesh, atmo, u_p, v_p, ps_p, t_p, δu, δv, δt, dt, tmp)
        correct(vc, mesh, atmo, mps, me, u_n, v_n, tmp, ps_n, spe_lnps_n, t_n, spe_t_n)

        wts = _gauss_weights(nlat)
        Δb = np.diff(vc.bk)
        # mass: the mean surface pressure is restored
        errs.append(abs(_global_mean(ps_n[0], wts) / _global_mean(ps_p[0], wts) - 1.0))

        def energy(u, v, t, ps):
            e = 0.5 * (u ** 2 + v ** 2) + atmo.cp_air * t
            col = (e * Δb[:, None, None]).sum(axis=0) * ps[0]
            return _global_mean(col, wts) / atmo.grav

        target = energy(u_p + dt * δu, v_p + dt * δv, t_p + dt * δt, ps_p)
        errs.append(abs(energy(u_n, v_n, t_n, ps_n) / target - 1.0))
        # the grid shifts are uniform constants, and the (0,0) coefficients
        # (P_00 = 1: a constant field c has coefficient c) shift by them
        # (differences of O(300) values: bounded on the scale of the shifted field)
        dT = t_n - t_n0
        errs.append(np.abs(dT - dT.flat[0]).max() / np.abs(t_n0).max())
        errs.append(np.abs((spe_t_n - spe_t_0)[:, 0, 0] - dT.flat[0]).max()
                    / max(1.0, np.abs(spe_t_0[:, 0, 0]).max()))
        rest = spe_t_n - spe_t_0
        rest[:, 0, 0] = 0
        errs.append(np.abs(rest).max())
        f = ps_n / ps_n0
        errs.append(np.abs(f - f.flat[0]).max())
        dl = (spe_lnps_n - spe_lnps_0)[0, 0, 0]
        errs.append(abs(dl - np.log(f.flat[0])) / abs(np.log(f.flat[0])))
    # the (0,0) convention itself: a constant grid field has (0,0) coefficient c
    c = 3.7
    s = mesh.trans_grid_to_spherical(np.full((1, nlat, nlon), c), np.zeros((1, N + 2, N + 1), complex))
    errs.append(abs(s[0, 0, 0] - c) / c)
    return max(errs)


def test_correction_pins():
    assert pin_corrections() < 1e-12


# ------------------------------------------------------------ planted mistakes
def _mutant(func, old, new):
    """A copy of ``func`` with one source substitution, compiled in the
    oracle module's namespace (CPU only)."""
    src = inspect.getsource(func)
    assert src.count(old) == 1, f"mutation site not found in {func.__name__}: {old!r}"
    ns = dict(vars(inspect.getmodule(func)))
    exec(compile(src.replace(old, new), f"<mutant of {func.__name__}>", "exec"), ns)
    f = ns[func.__name__]
    assert isinstance(f, types.FunctionType)
    return f


MUTANTS = {
    "a: sign of r (w[k+1] - w[k]) flipped in Vert_Advection!": (
        pin_vert_advection, P.vert_advection,
        "r * (w[1:nd + 1] - w[0:nd])", "r * (w[0:nd] - w[1:nd + 1])", None),
    "b: b_k and b_{k+1} swapped in Four_In_One!'s x1": (
        pin_four_in_one, P.four_in_one,
        "x1 = (bk[k] * Δlnp_m + bk[k + 1] * Δlnp_p) / dp",
        "x1 = (bk[k + 1] * Δlnp_m + bk[k] * Δlnp_p) / dp", None),
    "c: energy correction without g": (
        pin_corrections, P.compute_corrections,
        "corr = atmo.grav * (mean_energy_p - mean_energy_n)",
        "corr = (mean_energy_p - mean_energy_n)", "correct"),
    "d: energy target from u_p instead of u_p + dt du": (
        pin_corrections, P.compute_corrections_init,
        "(u_p + dt * δu) ** 2", "u_p ** 2", "init"),
}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistakes_fail_the_pins(name):
    pin, func, old, new, kw = MUTANTS[name]
    bad = _mutant(func, old, new)
    err = pin(**{kw: bad}) if kw else pin(bad)
    assert err > 1e-9, f"{name}: pin error {err:.2e} did not expose the mistake"
