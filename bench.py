"""Yee cell-update throughput of the 3D SOI 1x2 MMI workload on one MI355X.

Workload (default): 480 x 240 x 160 cells at 25 nm, six FP32 field components,
SOI stack with a seeded MMI, 12-cell PML, dipole Gaussian-pulse source, two
x-normal DFT flux ports (8 wavelengths). A step is one full Yee step (H and E
updates, source injection, port DFT accumulation) over the whole grid, with
all inputs resident in HBM.

Prints one JSON line: value = cell-updates/s; roofline = the fused Yee
kernel's algorithmic HBM bytes (56 B per cell-update: six FP32 components read
and written, FP32 permittivity and conductivity read) over its measured mean
launch time; cpu_baseline = the FP64 NumPy restatement of the reference's 3D
step (complex128 fields, as the reference) timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
B_ALG_3D = 56              # bytes per 3D cell-update
B_ALG_2D = 32


def cpu_baseline(nx, ny, nz, dx, dt, seconds_target=15.0):
    """FP64 NumPy restatement of fdtd.py:200-310 (complex128 fields), timed on 1 core."""
    from oracle import fdtd_ref as R
    rng = np.random.default_rng(0)
    eps = 1 + 10 * rng.random((nz, ny, nx)); sig = rng.random((nz, ny, nx)) * 1e3
    F = {k: np.zeros(s, dtype=np.complex128) for k, s in R.field_shapes(nx, ny, nz).items()}
    F["Ez"][nz // 2 - 1, ny // 2, nx // 2] = 1.0
    R.update_fields_3d(F, eps, sig, dx, dx, dx, dt)          # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        R.update_fields_3d(F, eps, sig, dx, dx, dx, dt)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_target or n >= 50:
            break
    return nx * ny * nz * n / el, n, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=480)
    ap.add_argument("--ny", type=int, default=240)
    ap.add_argument("--nz", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default=os.environ.get("BEAMZ_AMD_KERNEL3D", "fused"))
    ap.add_argument("--kseg", type=int, default=int(os.environ.get("BEAMZ_AMD_KSEG", "16")))
    ap.add_argument("--rows", type=int, default=int(os.environ.get("BEAMZ_AMD_ROWS", "4")))
    ap.add_argument("--timing-steps", type=int, default=30)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    import beamz_amd as bz
    from beamz_amd.workloads import soi_mmi_3d

    total = a.warmup + a.steps + a.timing_steps
    mesh, d3, t, mons = soi_mmi_3d(a.nx, a.ny, a.nz, steps=max(total, 16))
    sim = bz.FDTD(mesh, time=t)
    sim.grid.kernel = a.kernel
    sim.grid.kseg = a.kseg
    sim.grid.rows = a.rows
    sim.initialize_simulation(save=False, live=False)
    sim._advance(a.warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    sim._advance(a.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    cells = a.nx * a.ny * a.nz
    value = cells * a.steps / wall
    # kernel duration: HIP events around each update launch, on the launch stream, in a separate run
    sim.grid.start_timing()
    sim._advance(min(a.timing_steps, max(0, sim.num_steps - sim.current_step)))
    lt = sim.grid.launch_timer()
    sim.grid.timing = False
    step_kernel_s = lt if lt else gpu_s / a.steps
    achieved = cells * B_ALG_3D / step_kernel_s
    out = {
        "metric": "Yee cell-updates/s (3D six-component FP32 step)", "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D SOI 1x2 MMI Yee step {a.nx}x{a.ny}x{a.nz} @25nm, 12-cell PML, dipole pulse, "
                               f"2 DFT flux ports x 8 wavelengths", "cells": cells,
                   "kernel": sim.grid.kernel_name()},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel_ms": step_kernel_s * 1e3, "bytes_per_cell_update": B_ALG_3D},
    }
    if not a.no_cpu:
        dt = float(t[1] - t[0])
        v, n, el = cpu_baseline(a.nx, a.ny, max(4, a.nz // 8), 25e-9, dt)
        out["cpu_baseline"] = {"value": v, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                               "sample": f"{n} steps of the NumPy complex128 3D Yee update on "
                                         f"{a.nx}x{a.ny}x{max(4, a.nz // 8)} cells ({el:.1f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
