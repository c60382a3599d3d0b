"""Benchmark: SHT level-field transforms per second at T85 L20 on one MI355X.

Workload (one "step"): one forward (grid -> spectral) and one inverse
(spectral -> grid) transform of a batch of 61 level-fields at T85
(N=85, 256x128 Gaussian grid) -- the prognostic set of a T85 L20 Held-Suarez
state (vor 20, div 20, T 20, ln ps 1).  Inputs are synthetic, seeded,
band-limited fields resident in HBM before timing starts.  Throughput counts
each direction of each level-field as one transform.

The Legendre stage is measured in FP64 and in split-FP32; `value` is the
split-FP32 path when its measured error against the FP64 path is <= 1e-6,
otherwise the FP64 path.  The CPU baseline is the FP64 oracle restatement
(numpy, single thread) on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_TRUNC, NLAT, NLON, NFIELDS = 85, 128, 256, 61
HBM_PEAK = 8.0e12
BYTES_PER_FIELD = NLON * NLAT * 8 + (N_TRUNC + 1) * (N_TRUNC + 2) * 16  # one direction


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(seconds):
    from threadpoolctl import threadpool_limits
    from oracle import SpectralSphericalMesh
    with threadpool_limits(1):
        M = SpectralSphericalMesh(N_TRUNC, N_TRUNC + 1, NLAT, NLON, NFIELDS, 6371.0e3)
        g = np.random.default_rng(1).standard_normal((NFIELDS, NLAT, NLON))
        s = M.spec_zeros()
        reps = 0
        t0 = time.perf_counter()
        while True:
            M.trans_grid_to_spherical(g, s)
            M.trans_spherical_to_grid(s, g)
            reps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
    return {"value": 2 * NFIELDS * reps / el, "unit": "level-field transforms/s", "cores": 1,
            "kind": "port",
            "sample": f"{reps} forward+inverse passes of 61 T85 level-fields "
                      f"(numpy FP64 oracle, 1 thread, {el:.1f} s)"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)
    import torch
    import __graft_entry__
    jgcm = __graft_entry__.load_package()
    from jgcm_amd import _lib

    mesh = jgcm.Spectral_Spherical_Mesh(N_TRUNC, N_TRUNC + 1, NLAT, NLON, NFIELDS, 6371.0e3)
    rng = np.random.default_rng(2024)
    s0 = rng.standard_normal((NFIELDS, N_TRUNC + 2, N_TRUNC + 1)) + \
        1j * rng.standard_normal((NFIELDS, N_TRUNC + 2, N_TRUNC + 1))
    s0[:, :, 0] = s0[:, :, 0].real
    for m in range(N_TRUNC + 1):
        s0[:, :m, m] = 0
    s0[:, N_TRUNC + 1, :] = 0
    spec_in = torch.tensor(s0, device="cuda")
    grid = mesh.grid_zeros()
    spec = mesh.spectral_zeros()
    L = _lib.lib()
    h = mesh.handle
    stream = torch.cuda.ExternalStream(mesh.stream())

    def fwd(g, s):
        _lib.check(L.jgcm_trans_grid_to_spherical(h, ctypes.c_void_p(g.data_ptr()),
                                                  ctypes.c_void_p(s.data_ptr()), NFIELDS))

    def inv(s, g):
        _lib.check(L.jgcm_trans_spherical_to_grid(h, ctypes.c_void_p(s.data_ptr()),
                                                  ctypes.c_void_p(g.data_ptr()), NFIELDS))

    results = {}
    ref_grid = ref_spec = None
    for prec in ("fp64", "split32"):
        mesh.set_precision(prec)
        for _ in range(a.warmup):
            inv(spec_in, grid)
            fwd(grid, spec)
        mesh.synchronize()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        t_inv = t_fwd = 0.0
        t0 = time.perf_counter()
        ev[0].record(stream)
        for k in range(a.steps):
            inv(spec_in, grid)
            fwd(grid, spec)
        ev[1].record(stream)
        mesh.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        total_ms = ev[0].elapsed_time(ev[1])
        # per-direction launch durations (separately bracketed, same stream)
        ev[2].record(stream)
        for k in range(a.steps):
            inv(spec_in, grid)
        ev[3].record(stream)
        mesh.synchronize()
        t_inv = ev[2].elapsed_time(ev[3]) / a.steps
        e4 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e4[0].record(stream)
        for k in range(a.steps):
            fwd(grid, spec)
        e4[1].record(stream)
        mesh.synchronize()
        t_fwd = e4[0].elapsed_time(e4[1]) / a.steps
        g_out = grid.cpu().numpy()
        s_out = spec.cpu().numpy()
        if prec == "fp64":
            ref_grid, ref_spec = g_out, s_out
            err = 0.0
        else:
            eg = (np.linalg.norm((g_out - ref_grid).reshape(NFIELDS, -1), axis=1) /
                  np.linalg.norm(ref_grid.reshape(NFIELDS, -1), axis=1)).max()
            es = (np.linalg.norm((s_out - ref_spec).reshape(NFIELDS, -1), axis=1) /
                  np.linalg.norm(ref_spec.reshape(NFIELDS, -1), axis=1)).max()
            err = float(max(eg, es))
        ms_step = total_ms / a.steps
        results[prec] = {"ms_per_step": ms_step, "fields_per_s": 2 * NFIELDS / (ms_step * 1e-3),
                         "ms_inverse": t_inv, "ms_forward": t_fwd, "wall_s": wall,
                         "max_rel_err_vs_fp64": err}
    mesh.set_precision("fp64")
    use = "split32" if results["split32"]["max_rel_err_vs_fp64"] <= 1e-6 else "fp64"
    r = results[use]
    # roofline of the dominant launch (the slower direction), algorithmic bytes
    dom = "inverse" if r["ms_inverse"] >= r["ms_forward"] else "forward"
    t_dom = max(r["ms_inverse"], r["ms_forward"]) * 1e-3
    achieved = NFIELDS * BYTES_PER_FIELD / t_dom
    out = {
        "metric": "SHT level-field transforms/s at T85 (61 level-fields per launch), "
                  "split-FP32 vs FP64 Legendre at <=1e-6 rel err",
        "value": r["fields_per_s"], "unit": "level-field transforms/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-split(f64 I/O)" if use == "split32" else "f64",
        "data": "synthetic",
        "config": {"workload": "T85 (N=85, 256x128) forward+inverse SHT of 61 level-fields",
                   "path": use},
        "precisions": results,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": f"{dom} transform ({use})"},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
