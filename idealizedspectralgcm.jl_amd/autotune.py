"""Autotuner for the launch shapes the library leaves open at run time, with a
persistent cache.

Two shapes are tuned, separately for the forward transforms (grid ->
spectral) and the inverse transforms (spectral -> grid):

* the longitude FFT's workgroup shape (latitude pairs per workgroup,
  ``jgcm_mesh_set_fft_shape``) of the fold (forward) and unfold (inverse);
* the Legendre stage's tile shape (16-row result tiles per workgroup of the
  forward quadrature, 16-latitude tiles of the inverse synthesis,
  ``jgcm_mesh_set_legendre_shape``).

They depend on the device, the grid and truncation, the Legendre precision
and the number of level-fields per launch.  The result of a transform does
not depend on either shape, only its time does.  ``Tune_Shapes`` times the
FFT candidates with the built-in Legendre tiles, then the tile candidates
with the chosen FFT shapes (coordinate descent: 5 + 3 timings per direction
instead of 15).

Cache format (JSON, one file; ``JGCM_TUNING_FILE`` or
``~/.cache/jgcm_amd/tuning.json``)::

    {"version": 2,
     "entries": {"<device>": {"T85|128x256|split32|61": {
         "fold_ppw": 4, "unfold_ppw": 8, "fwd_tiles": 2, "inv_tiles": 1,
         "forward_us": {"1": 31.2, "2": 28.0, ...},
         "inverse_us": {"1": 30.5, ...},
         "legendre_forward_us": {"1": 27.9, "2": 26.1, "4": 26.4},
         "legendre_inverse_us": {"1": 25.0, ...}}}}}

``<device>`` is the device name and architecture as torch reports them;
``*_us`` are the timings of whole transforms per candidate.  The reference
has no counterpart (FFTW plans per call on the CPU, and its Legendre sums are
plain loops).
"""
import ctypes
import json
import os

from ._lib import call

CANDIDATES = (1, 2, 4, 8, 16)
TILE_CANDIDATES = (1, 2, 4)
VERSION = 2


def default_cache_path():
    p = os.environ.get("JGCM_TUNING_FILE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "jgcm_amd", "tuning.json")


def device_key(index=None):
    """'<name>|<arch>' of a CUDA/HIP device as torch reports it."""
    import torch
    i = torch.cuda.current_device() if index is None else index
    prop = torch.cuda.get_device_properties(i)
    return f"{prop.name}|{getattr(prop, 'gcnArchName', '')}"


def tuning_key(mesh, nfields):
    return f"T{mesh.num_fourier}|{mesh.nθ}x{mesh.nλ}|{mesh.precision}|{int(nfields)}"


class Tuning_Cache:
    """The JSON cache: entries[device][key] -> dict.  Writes go through a
    temporary file and a rename, so a reader never sees half a file."""

    def __init__(self, path=None):
        self.path = path or default_cache_path()
        self.data = {"version": VERSION, "entries": {}}
        if os.path.exists(self.path):
            with open(self.path) as f:
                d = json.load(f)
            if d.get("version") == VERSION and isinstance(d.get("entries"), dict):
                self.data = d

    def get(self, device, key):
        return self.data["entries"].get(device, {}).get(key)

    def put(self, device, key, entry):
        self.data["entries"].setdefault(device, {})[key] = entry
        self.save()

    def save(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def get_fft_shape(mesh):
    """(fold_ppw, unfold_ppw, max_ppw) of a mesh; 0 = the built-in rule."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("jgcm_mesh_get_fft_shape", mesh.handle, ctypes.byref(a), ctypes.byref(b),
         ctypes.byref(c))
    return a.value, b.value, c.value


def set_fft_shape(mesh, fold_ppw, unfold_ppw):
    call("jgcm_mesh_set_fft_shape", mesh.handle, int(fold_ppw), int(unfold_ppw))


def get_legendre_shape(mesh):
    """(forward tiles, inverse tiles) of a mesh; 0 = the built-in 2."""
    a, b = ctypes.c_int(), ctypes.c_int()
    call("jgcm_mesh_get_legendre_shape", mesh.handle, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def set_legendre_shape(mesh, fwd_tiles, inv_tiles):
    call("jgcm_mesh_set_legendre_shape", mesh.handle, int(fwd_tiles), int(inv_tiles))


def select(forward_us, inverse_us):
    """Fastest candidate per direction; ties go to the smaller shape."""
    fold = min(forward_us, key=lambda c: (forward_us[c], c))
    unfold = min(inverse_us, key=lambda c: (inverse_us[c], c))
    return int(fold), int(unfold)


def _time_us(mesh, fn, reps):
    import torch
    stream = torch.cuda.ExternalStream(mesh.stream())
    for _ in range(3):
        fn()
    mesh.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    mesh.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def _inputs(mesh, nfields, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    grid = torch.randn((nfields, mesh.nθ, mesh.nλ), generator=g, dtype=torch.float64).to(mesh.device)
    spec = mesh.spectral_zeros(nfields)
    torch.cuda.synchronize()
    return grid, spec


def _sweep(mesh, cands, apply, grid, spec, reps, workload=None):
    """Forward and inverse times (us) per candidate: of plain transforms of
    `grid` / `spec`, or of the workload's two callables."""
    from .mesh import Trans_Grid_To_Spherical, Trans_Spherical_To_Grid
    if workload is not None:
        _, run_fwd, run_inv = workload
    else:
        run_fwd = lambda: Trans_Grid_To_Spherical(mesh, grid, spec, batch=True)
        run_inv = lambda: Trans_Spherical_To_Grid(mesh, spec, grid, batch=True)
    fwd, inv = {}, {}
    for c in cands:
        apply(c)
        fwd[c] = _time_us(mesh, run_fwd, reps)
        inv[c] = _time_us(mesh, run_inv, reps)
    return fwd, inv


def _key(mesh, nfields, workload):
    k = tuning_key(mesh, nfields)
    return k if workload is None else f"{k}|{workload[0]}"


def _us(d):
    return {str(c): round(v, 3) for c, v in d.items()}


def Tune_FFT_Shape(mesh, nfields, candidates=CANDIDATES, reps=30, cache=None, seed=0,
                   workload=None):
    """Time forward and inverse transforms of `nfields` synthetic level-fields
    on `mesh` for every candidate FFT shape (those above the mesh's LDS limit
    are skipped; the mesh's Legendre tiles stay as they are), apply the
    fastest fold and unfold shapes and return the entry; stored in `cache`
    (a Tuning_Cache) when given.  `workload` = (name, forward, inverse):
    time those two callables (a model's own transform batches) instead; the
    cache key then carries the name."""
    limit = get_fft_shape(mesh)[2]
    cands = [c for c in candidates if 1 <= c <= limit]
    if not cands:
        raise ValueError(f"no candidate shape fits (limit {limit})")
    grid, spec = _inputs(mesh, nfields, seed) if workload is None else (None, None)
    fwd, inv = _sweep(mesh, cands, lambda c: set_fft_shape(mesh, c, c), grid, spec, reps,
                      workload)
    fold, unfold = select(fwd, inv)
    set_fft_shape(mesh, fold, unfold)
    ft, it = get_legendre_shape(mesh)
    entry = {"fold_ppw": fold, "unfold_ppw": unfold, "fwd_tiles": ft, "inv_tiles": it,
             "forward_us": _us(fwd), "inverse_us": _us(inv)}
    if cache is not None:
        cache.put(device_key(), _key(mesh, nfields, workload), entry)
    return entry


def Tune_Legendre_Shape(mesh, nfields, candidates=TILE_CANDIDATES, reps=30, seed=0,
                        workload=None):
    """Time forward and inverse transforms for every Legendre tile candidate
    (the mesh's FFT shapes stay as they are), apply the fastest per
    direction and return (fwd_tiles, inv_tiles, forward_us, inverse_us)."""
    cands = [c for c in candidates if c in TILE_CANDIDATES]
    if not cands:
        raise ValueError(f"Legendre tile candidates must be among {TILE_CANDIDATES}")
    grid, spec = _inputs(mesh, nfields, seed) if workload is None else (None, None)
    fwd, inv = _sweep(mesh, cands, lambda c: set_legendre_shape(mesh, c, c), grid, spec, reps,
                      workload)
    ft, it = select(fwd, inv)
    set_legendre_shape(mesh, ft, it)
    return ft, it, fwd, inv


def Tune_Shapes(mesh, nfields, reps=30, cache=None, seed=0, workload=None):
    """FFT shapes (with the built-in Legendre tiles), then Legendre tiles
    (with the chosen FFT shapes); both applied to `mesh`, the entry returned
    and stored in `cache` when given."""
    set_legendre_shape(mesh, 0, 0)
    entry = Tune_FFT_Shape(mesh, nfields, reps=reps, seed=seed, workload=workload)
    ft, it, lf, li = Tune_Legendre_Shape(mesh, nfields, reps=reps, seed=seed, workload=workload)
    entry.update({"fwd_tiles": ft, "inv_tiles": it,
                  "legendre_forward_us": _us(lf), "legendre_inverse_us": _us(li)})
    if cache is not None:
        cache.put(device_key(), _key(mesh, nfields, workload), entry)
    return entry


def Apply_Tuning(mesh, nfields, cache=None, tune_missing=False, workload=None, **kw):
    """Set the mesh's FFT and Legendre shapes from the cache entry for
    (device, mesh, nfields[, workload name]); with `tune_missing`, tune and
    store when there is none.  Returns the entry used, or None (the built-in
    shapes stay)."""
    cache = cache if cache is not None else Tuning_Cache()
    entry = cache.get(device_key(), _key(mesh, nfields, workload))
    if entry is None:
        if not tune_missing:
            return None
        return Tune_Shapes(mesh, nfields, cache=cache, workload=workload, **kw)
    limit = get_fft_shape(mesh)[2]
    fold, unfold = int(entry["fold_ppw"]), int(entry["unfold_ppw"])
    ft, it = int(entry.get("fwd_tiles", 0)), int(entry.get("inv_tiles", 0))
    tiles_ok = all(t == 0 or t in TILE_CANDIDATES for t in (ft, it))
    if not (1 <= fold <= limit and 1 <= unfold <= limit and tiles_ok):
        return None
    set_fft_shape(mesh, fold, unfold)
    set_legendre_shape(mesh, ft, it)
    return entry
