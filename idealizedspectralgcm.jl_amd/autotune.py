"""Autotuner for the launch shapes the library leaves open at run time, with a
persistent cache.

The shape tuned is the longitude FFT's workgroup shape (latitude pairs per
workgroup, ``jgcm_mesh_set_fft_shape``), separately for the fold of the
forward transforms and the unfold of the inverse transforms.  It depends on
the device, the grid and truncation, the Legendre precision (through the
kernels running beside it) and the number of level-fields per launch.  The
result of a transform does not depend on the shape, only its time does.

Cache format (JSON, one file; ``JGCM_TUNING_FILE`` or
``~/.cache/jgcm_amd/tuning.json``)::

    {"version": 1,
     "entries": {"<device>": {"T85|128x256|split32|61": {
         "fold_ppw": 4, "unfold_ppw": 8,
         "forward_us": {"1": 31.2, "2": 28.0, ...},
         "inverse_us": {"1": 30.5, ...}}}}}

``<device>`` is the device name and architecture as torch reports them.  The
reference has no counterpart (FFTW plans per call on the CPU).
"""
import ctypes
import json
import os

from ._lib import call

CANDIDATES = (1, 2, 4, 8, 16)
VERSION = 1


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


def Tune_FFT_Shape(mesh, nfields, candidates=CANDIDATES, reps=30, cache=None, seed=0):
    """Time forward and inverse transforms of `nfields` synthetic level-fields
    on `mesh` for every candidate shape (those above the mesh's LDS limit are
    skipped), apply the fastest fold and unfold shapes and return the entry;
    stored in `cache` (a Tuning_Cache) when given."""
    import torch
    from .mesh import Trans_Grid_To_Spherical, Trans_Spherical_To_Grid
    limit = get_fft_shape(mesh)[2]
    cands = [c for c in candidates if 1 <= c <= limit]
    if not cands:
        raise ValueError(f"no candidate shape fits (limit {limit})")
    g = torch.Generator(device="cpu").manual_seed(seed)
    grid = torch.randn((nfields, mesh.nθ, mesh.nλ), generator=g, dtype=torch.float64).to(mesh.device)
    spec = mesh.spectral_zeros(nfields)
    torch.cuda.synchronize()
    fwd, inv = {}, {}
    for c in cands:
        set_fft_shape(mesh, c, c)
        fwd[c] = _time_us(mesh, lambda: Trans_Grid_To_Spherical(mesh, grid, spec, batch=True), reps)
        inv[c] = _time_us(mesh, lambda: Trans_Spherical_To_Grid(mesh, spec, grid, batch=True), reps)
    fold, unfold = select(fwd, inv)
    set_fft_shape(mesh, fold, unfold)
    entry = {"fold_ppw": fold, "unfold_ppw": unfold,
             "forward_us": {str(c): round(v, 3) for c, v in fwd.items()},
             "inverse_us": {str(c): round(v, 3) for c, v in inv.items()}}
    if cache is not None:
        cache.put(device_key(), tuning_key(mesh, nfields), entry)
    return entry


def Apply_Tuning(mesh, nfields, cache=None, tune_missing=False, **kw):
    """Set the mesh's FFT shape from the cache entry for (device, mesh,
    nfields); with `tune_missing`, tune and store when there is none.
    Returns the entry used, or None (the built-in rule stays)."""
    cache = cache if cache is not None else Tuning_Cache()
    entry = cache.get(device_key(), tuning_key(mesh, nfields))
    if entry is None:
        if not tune_missing:
            return None
        return Tune_FFT_Shape(mesh, nfields, cache=cache, **kw)
    limit = get_fft_shape(mesh)[2]
    fold, unfold = int(entry["fold_ppw"]), int(entry["unfold_ppw"])
    if not (1 <= fold <= limit and 1 <= unfold <= limit):
        return None
    set_fft_shape(mesh, fold, unfold)
    return entry
