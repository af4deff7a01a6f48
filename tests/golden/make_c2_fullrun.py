"""Writes tests/golden/c2_fullrun.npz: the FP64 oracle's result of the full 3D workload (480 x 240 x 160 cells,
dx = 25 nm, seeded SOI 1x2 MMI, 12-cell PML, dipole Gaussian pulse, 3000 steps, two x-normal flux ports at
8 wavelengths), too long for the oracle inside a GPU test (about 7 minutes on 8 cores).

Stored (complex64 / float32 carry the FP64 values to 6e-8, far below the 1e-4 tolerance):
- dft_p{q}_{Ey,Ez,Hy,Hz}: each port plane's DFT (nfreq, z, y) of the components the x-normal flux uses;
- flux_p{q}: the net flux spectrum of each port (FP64);
- sample_{c}: the final field of each component at SAMPLE seeded cells (rng seed SEED, chosen without
  replacement over the component's flat reference-shape array), and norm_{c} the full array's L2 norm;
- floor_{c}: rel-L2 of the same run with fields stored in FP32 (oracle step_3d_c_f32store) against FP64,
  over the whole array: the accuracy floor of FP32 field storage on this run.

Run from the repository root: python tests/golden/make_c2_fullrun.py (needs oracle/liboref.so and
beamz_amd/libbeamz_amd.so for the rasterizer)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from beamz_amd.workloads import soi_mmi_3d  # noqa: E402
from oracle import fdtd_ref as R  # noqa: E402

SEED = 2026
SAMPLE = 8192
COMP3 = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_fullrun.npz")


def sample_index(shape, comp):
    rng = np.random.default_rng([SEED, COMP3.index(comp)])
    return np.sort(rng.choice(int(np.prod(shape)), SAMPLE, replace=False))


def main():
    t0 = time.time()
    mesh, d3, t, mons = soi_mmi_3d(480, 240, 160, steps=3000, record_ports=True)
    dt = float(t[1] - t[0])
    kw = dict(dz=mesh.dz, sources=d3.sources, design_dims=(d3.width, d3.height, d3.depth), native=True,
              true_planes=True)
    ref = R.RefSim(mesh.permittivity, mesh.conductivity, mesh.dx, mesh.dy, dt, len(t), monitors=mons, **kw)
    for c in range(0, len(t), 500):
        ref.run(500)
        print(f"FP64 oracle: {c + 500} steps, {time.time() - t0:.0f} s", flush=True)
    out = {}
    for q, m in enumerate(mons):
        ts = ref.mon_series[q]["t"]
        D = {k: R.dft_of_series(np.array(ref.mon_series[q][k]), ts, 2 * np.pi * m.dft_frequencies)
             for k in ("Ey", "Ez", "Hy", "Hz")}
        for k, v in D.items():
            out[f"dft_p{q + 1}_{k}"] = v.astype(np.complex64)
        out[f"flux_p{q + 1}"] = R.plane_flux(D, "x")
    ref.mon_series = None
    final = {k: v.copy() for k, v in ref.F.items()}
    del ref
    for k in COMP3:
        out[f"sample_{k}"] = final[k].ravel()[sample_index(final[k].shape, k)].astype(np.float32)
        out[f"norm_{k}"] = np.float64(np.linalg.norm(final[k]))
    flo = R.RefSim(mesh.permittivity, mesh.conductivity, mesh.dx, mesh.dy, dt, len(t), f32_storage=True, **kw)
    for c in range(0, len(t), 500):
        flo.run(500)
        print(f"FP32-storage floor: {c + 500} steps, {time.time() - t0:.0f} s", flush=True)
    for k in COMP3:
        out[f"floor_{k}"] = np.float64(np.linalg.norm(flo.F[k] - final[k]) / np.linalg.norm(final[k]))
    out["seed"], out["sample"], out["steps"] = np.int64(SEED), np.int64(SAMPLE), np.int64(len(t))
    np.savez(OUT, **out)
    print({k: f"{float(out['floor_' + k]):.2e}" for k in COMP3}, f"{os.path.getsize(OUT) / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
