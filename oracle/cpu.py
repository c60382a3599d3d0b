"""ctypes binding and build recipe of the C++ FP64 transform restatement
(``oracle/sht_cpu.cpp``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): bench.py times it as
the CPU baseline on the host cores, tests/test_oracle_cpu.py checks it
against the numpy oracle.  ``build()`` compiles two variants with g++ into
``oracle/_build/`` (x86-64-v3 with AVX2/FMA, and baseline x86-64);
``load()`` takes the first one the host CPU can run.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "sht_cpu.cpp")
BUILD = os.path.join(HERE, "_build")
VARIANTS = (("v3", ["-march=x86-64-v3"]), ("base", ["-march=x86-64"]))
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fno-fast-math"]


def _path(tag):
    return os.path.join(BUILD, f"libsht_cpu_{tag}.so")


def build(verbose=False):
    """Compile every variant (g++; about 5 s)."""
    os.makedirs(BUILD, exist_ok=True)
    for tag, arch in VARIANTS:
        out = _path(tag)
        cmd = ["g++", *FLAGS, *arch, SRC, "-o", out + ".tmp"]
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
            lib.variant = tag
            _LIB = lib
            return lib
    raise RuntimeError("the C++ transform restatement is not built (oracle.cpu.build())")


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
