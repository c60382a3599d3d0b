"""ctypes binding and build recipe of the C++ FP64 restatements: the
transform pair (``oracle/sht_cpu.cpp``) and the Held-Suarez step
(``oracle/hs_cpu.cpp``), one library.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): bench.py times them as
the CPU baselines on the host cores, tests/test_oracle_cpu.py checks them
against the numpy oracle.  ``build()`` compiles two variants with g++ into
``oracle/_build/`` (x86-64-v3 with AVX2/FMA, and baseline x86-64);
``load()`` takes the first one the host CPU can run.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, f) for f in ("sht_cpu.cpp", "hs_cpu.cpp")]
HEADERS = [os.path.join(HERE, "fft_cpu.hpp")]
BUILD = os.path.join(HERE, "_build")
VARIANTS = (("v3", ["-march=x86-64-v3"]), ("base", ["-march=x86-64"]))
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fno-fast-math"]


def _path(tag):
    return os.path.join(BUILD, f"libjgcm_oracle_cpu_{tag}.so")


def build(verbose=False):
    """Compile every variant (g++; about 5 s)."""
    os.makedirs(BUILD, exist_ok=True)
    for tag, arch in VARIANTS:
        out = _path(tag)
        cmd = ["g++", *FLAGS, *arch, *SRC, "-o", out + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed:\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)


def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


_LIB = None


def load():
    """The fastest variant the host runs (raises if none is built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    flags = _cpu_flags()
    v3_ok = {"avx2", "fma", "bmi2"} <= flags
    for tag, _ in VARIANTS:
        p = _path(tag)
        if tag == "v3" and not v3_ok:
            continue
        if os.path.exists(p):
            lib = ctypes.CDLL(p)
            P, I = ctypes.c_void_p, ctypes.c_int
            for name in ("sht_forward", "sht_inverse"):
                fn = getattr(lib, name)
                fn.argtypes = [I, I, I, I, I, P, P, P, I]
                fn.restype = I
            lib.sht_max_threads.restype = I
            lib.hs_step.argtypes = [P, P, I]
            lib.hs_step.restype = I
            lib.hs_config_size.restype = I
            lib.hs_state_size.restype = I
            lib.variant = tag
            _LIB = lib
            return lib
    raise RuntimeError("the C++ restatements are not built (oracle.cpu.build())")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class CpuSHT:
    """The transform pair of one oracle mesh (tables taken from it)."""

    def __init__(self, mesh):
        self.mesh = mesh
        N, ns, nlat = mesh.num_fourier, mesh.num_spherical, mesh.nlat
        self.qnm = np.ascontiguousarray(mesh.qnm[:, :ns + 1, :], dtype=np.float64)
        self.qwg = np.ascontiguousarray(mesh.qwg[:, :ns + 1, :], dtype=np.float64)
        assert self.qnm.shape == (N + 1, ns + 1, nlat)
        self.lib = load()

    def forward(self, grid, snm, threads=1):
        """Trans_Grid_To_Spherical! (SSM:221-293): grid [nlev, nθ, nλ] ->
        snm complex [nlev, ns+1, N+1] (lower triangle untouched)."""
        M = self.mesh
        assert grid.dtype == np.float64 and grid.flags.c_contiguous
        assert snm.dtype == np.complex128 and snm.flags.c_contiguous
        rc = self.lib.sht_forward(M.num_fourier, M.num_spherical, M.nlat, M.nlon, grid.shape[0],
                                  _p(self.qwg), _p(grid), _p(snm), int(threads))
        assert rc == 0
        return snm

    def inverse(self, snm, grid, threads=1):
        """Trans_Spherical_To_Grid! (SSM:143-217)."""
        M = self.mesh
        assert grid.dtype == np.float64 and grid.flags.c_contiguous
        assert snm.dtype == np.complex128 and snm.flags.c_contiguous
        rc = self.lib.sht_inverse(M.num_fourier, M.num_spherical, M.nlat, M.nlon, snm.shape[0],
                                  _p(self.qnm), _p(snm), _p(grid), int(threads))
        assert rc == 0
        return grid


# hs_cpu.cpp HsState, in its field order: (C field, oracle DynData attribute)
HS_STATE = [(a.replace("δ", "d").replace("Δ", "D").replace("λ", "lam").replace("θ", "th"), a)
            for a in (
    "grid_u_p", "grid_v_p", "grid_t_p", "grid_u_c", "grid_v_c", "grid_t_c",
    "grid_u_n", "grid_v_n", "grid_t_n", "grid_δu", "grid_δv", "grid_δt", "grid_t_eq",
    "grid_p_full", "grid_lnp_full", "grid_Δp", "grid_w_full", "grid_geopot_full",
    "grid_energy_full", "grid_d_full1", "grid_vor", "grid_div", "grid_absvor", "grid_z_full",
    "grid_p_half", "grid_lnp_half", "grid_M_half", "grid_geopot_half", "grid_z_half",
    "grid_ps_p", "grid_ps_c", "grid_ps_n", "grid_δps", "grid_δlnps", "grid_lnps",
    "grid_dλ_ps", "grid_dθ_ps", "grid_geopots",
    "spe_vor_p", "spe_vor_c", "spe_vor_n", "spe_div_p", "spe_div_c", "spe_div_n",
    "spe_t_p", "spe_t_c", "spe_t_n", "spe_δvor", "spe_δdiv", "spe_δt", "spe_energy",
    "spe_lnps_p", "spe_lnps_c", "spe_lnps_n", "spe_δlnps")]

_I, _D, _P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_CONFIG_INTS = ("N", "ns", "nlat", "nlon", "nd", "num_wavenumbers", "zero_top", "centered",
                "do_mass", "do_energy", "init_step")
_CONFIG_DOUBLES = ("radius", "grav", "rdgas", "cp_air", "kappa", "dt", "robert", "implicit_coef",
                   "ps_ref", "sigma_b", "k_f", "k_a", "k_s", "dT_y", "dth_z")
_CONFIG_TABLES = ("ak", "bk", "qnm", "qwg", "alp_a", "alp_b", "uvc", "uvm", "uvp", "dlam", "dthm",
                  "dthp", "eig", "damping", "wts", "cos_theta", "sin_theta", "coriolis", "dp_ref",
                  "lnp_half_ref", "lnp_full_ref", "t_ref", "h", "wave_matrix")


class _HsConfig(ctypes.Structure):
    _fields_ = ([(n, _I) for n in _CONFIG_INTS] + [(n, _D) for n in _CONFIG_DOUBLES]
                + [(n, _P) for n in _CONFIG_TABLES])


class _HsState(ctypes.Structure):
    _fields_ = [(n, _P) for n, _ in HS_STATE]


class CpuHS:
    """Atmosphere_Update! (Spectral_Dynamics.jl:470-490) of the numpy oracle's
    objects -- mesh, atmosphere, vertical coordinate, semi-implicit solver (its
    integrator and W_n matrices), Dyn_Data and Held-Suarez parameters, the
    arguments of ``primitive.atmosphere_update`` in its order -- in
    C++ (oracle/hs_cpu.cpp), updating the Dyn_Data arrays in place exactly
    where ``primitive.atmosphere_update`` does."""

    def __init__(self, mesh, atmo, vc, si, d, params):
        lib = load()
        assert lib.hs_config_size() == ctypes.sizeof(_HsConfig)
        assert lib.hs_state_size() == ctypes.sizeof(_HsState)
        assert vc.vert_advect_scheme in ("second_centered_wts", "second_centered")
        M, ns = mesh, mesh.num_spherical
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.tables = {
            "ak": f64(vc.ak), "bk": f64(vc.bk),
            "qnm": f64(M.qnm[:, :ns + 1, :]), "qwg": f64(M.qwg[:, :ns + 1, :]),
            "alp_a": f64(M.coef_alp_a), "alp_b": f64(M.coef_alp_b),
            "uvc": f64(M.coef_uvc), "uvm": f64(M.coef_uvm), "uvp": f64(M.coef_uvp),
            "dlam": f64(M.coef_dlam), "dthm": f64(M.coef_dthm), "dthp": f64(M.coef_dthp),
            "eig": f64(M.laplacian_eig), "wts": f64(M.wts), "cos_theta": f64(M.cos_theta),
            "sin_theta": f64(M.sin_theta), "coriolis": f64(atmo.coriolis),
            "dp_ref": f64(si.Δp_ref), "lnp_half_ref": f64(si.lnp_half_ref),
            "lnp_full_ref": f64(si.lnp_full_ref), "t_ref": f64(si.t_ref), "h": f64(si.h)}
        for a in self.tables.values():
            assert a.flags.c_contiguous
        self.mesh, self.vc, self.atmo, self.si, self.d, self.params = M, vc, atmo, si, d, params
        self.lib = lib
        st = _HsState()
        for cname, attr in HS_STATE:
            a = getattr(d, attr)
            assert a.flags.c_contiguous and a.dtype in (np.float64, np.complex128), attr
            setattr(st, cname, a.ctypes.data)
        self.state = st

    def _config(self):
        M, vc, atmo, si, p = self.mesh, self.vc, self.atmo, self.si, self.params
        I = si.integrator
        self.tables["damping"] = np.ascontiguousarray(I.damping, dtype=np.float64)
        self.tables["wave_matrix"] = np.ascontiguousarray(si.wave_matrix, dtype=np.float64)
        c = _HsConfig()
        ints = dict(N=M.num_fourier, ns=M.num_spherical, nlat=M.nlat, nlon=M.nlon, nd=vc.nd,
                    num_wavenumbers=si.num_wavenumbers, zero_top=int(vc.zero_top),
                    centered=int(vc.vert_advect_scheme == "second_centered"),
                    do_mass=int(atmo.do_mass_correction), do_energy=int(atmo.do_energy_correction),
                    init_step=int(I.init_step))
        dbl = dict(radius=M.radius, grav=atmo.grav, rdgas=atmo.rdgas, cp_air=atmo.cp_air,
                   kappa=atmo.kappa, dt=I.dt, robert=I.robert_coef, implicit_coef=I.implicit_coef,
                   ps_ref=si.ps_ref, sigma_b=p["σ_b"], k_f=p["k_f"], k_a=p["k_a"], k_s=p["k_s"],
                   dT_y=p["ΔT_y"], dth_z=p["Δθ_z"])
        for k, v in ints.items():
            setattr(c, k, int(v))
        for k, v in dbl.items():
            setattr(c, k, float(v))
        for k in _CONFIG_TABLES:
            setattr(c, k, self.tables[k].ctypes.data)
        return c

    def step(self, threads=1):
        """One Atmosphere_Update! on ``threads`` OpenMP threads."""
        assert not self.atmo.do_water_correction
        c = self._config()
        rc = self.lib.hs_step(ctypes.byref(c), ctypes.byref(self.state), int(threads))
        if rc == -2:
            raise AssertionError("HS_Forcing: a full level lies below the surface (σ > 1)")
        assert rc == 0, rc
