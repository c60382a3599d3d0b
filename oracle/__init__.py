"""FP64 CPU oracle for the JGCM spectral-transform core.

TEST INFRASTRUCTURE ONLY.  This package is a numpy restatement of the
reference Julia code (CliMA/IdealizedSpectralGCM.jl, ``src/Atmos_Spectral``
and ``src/Atmos_Param``).  It exists to check the HIP library: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``libjgcm_hip.so`` and the
host mirror in ``idealizedspectralgcm.jl_amd``) never calls into it.
``cpu.py`` binds C++ restatements of the transform pair and the Held-Suarez
step (``sht_cpu.cpp``, ``hs_cpu.cpp``), checked against the numpy functions
and timed by bench.py as the CPU baselines.

Array layout is the reference's memory layout read in C order:

* spectral fields: ``complex128[nd, num_spherical+1, num_fourier+1]``
  (Julia ``ComplexF64[N+1, N+2, nd]`` column-major, m fastest);
* grid fields: ``float64[nd, nθ, nλ]`` (Julia ``Float64[nλ, nθ, nd]``),
  latitude index j runs south to north.

Every function cites the reference file:line it restates.  Parity pins
(the checks that make the restatement trustworthy) live in
``tests/test_oracle_*.py``; functions without such a check are listed as
"parity unpinned" in the module docstrings where they occur.
"""
from .gauss_legendre import compute_gaussian, compute_legendre
from .mesh import SpectralSphericalMesh
from .time_integrator import FilteredLeapfrog

__all__ = ["compute_gaussian", "compute_legendre", "SpectralSphericalMesh",
           "FilteredLeapfrog"]
