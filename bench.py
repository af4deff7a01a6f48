"""Yee cell-update throughput on one MI355X.

Default workload (the headline): the 3D SOI 1x2 MMI, 480 x 240 x 160 cells at
25 nm, six FP32 field components, seeded MMI, 12-cell PML, dipole
Gaussian-pulse source, two x-normal DFT flux ports (8 wavelengths). A step is
one full Yee step (H and E updates, source injection, port DFT accumulation)
over the whole grid, with all inputs resident in HBM.

``--workload wg2d`` runs the 2D TM straight-waveguide case (800 x 400 cells,
20-cell PML, point source, 16-wavelength DFT line monitor); ``--workload
batch2d`` the batch of 64 independent 2D simulations (sides drawn from
256..1024 cells by seed, binary 2 x 2 um design regions, 10-cell PML, input and
output DFT ports at 8 wavelengths), all advanced together, two steps per launch.

``--workload surrogate_train`` times the surrogate's BF16-MFMA training step
alone (synthetic binary 128 x 128 densities and targets, batch 128) and reports
FLOP/s against the dense BF16 MFMA peak.

Prints one JSON line: value = cell-updates/s (for the batch: summed over its
members); roofline = the update kernel's algorithmic HBM bytes (3D: 56 B per
cell-update, six FP32 components read and written plus FP32 permittivity and
conductivity read; 2D: 32 B) over its mean launch time measured with HIP
events on the launch stream; cpu_baseline = the FP64 NumPy restatement of the
reference's step (complex128 fields, as the reference computes) timed on a
bounded sample on one host core.
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
B_ALG = {"mmi3d": 56, "wg2d": 32, "batch2d": 32, "adjoint2d": 32, "surrogate": 32, "surrogate_train": 0}
MFMA_BF16_PEAK = 2.5e15    # MI355X dense BF16 MFMA, FLOP/s
INFINITY_CACHE = 256 * 2 ** 20    # MI355X last-level (MALL) cache, bytes

# HBM bytes per launch of the update kernel, from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of
# this script with --copy-mib 512, each counter scaled by the factor that makes the 512 MiB calibration
# copy read its known size (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0). Keyed by (workload, kernel, rows, kseg,
# nx, ny, nz); a run with any other configuration reports traffic as null. Source: the counter summary
# profiles/yee3d_fused_mmi_480x240x160_hbm_counters.md.
MEASURED_TRAFFIC = {
    ("mmi3d", "fused", 2, 16, 480, 240, 160): (874_454 + 476_068) * 1024,
    # two-step kernel: one launch = two steps (profiles/yee3d_tb2_mmi_480x240x160.md)
    ("mmi3d", "fused-tb2-w8", 2, 54, 480, 240, 160): (2 * 352_009 + 465_912) * 1024,
    # batched 2D engine, 64 members, 2-row tiles in row-major order
    # (profiles/yee2d_fused_batch64_kernel_and_counters.md)
    ("batch2d", "seg2", 64): (2 * 483_330 + 367_632) * 1024,
    # two steps per launch, 60 x 8 pair tiles in row-major order (profiles/yee2d_tb2_batch64_kernel_and_counters.md)
    ("batch2d", "pairs-seg8", 64): (2 * 523_987 + 403_190) * 1024,
    # the same in XCD order with the monitor gather of the previous pair in the same launch, the default
    # (same profile)
    ("batch2d", "pairs-seg8-xcd", 64): (2 * 474_825 + 400_315) * 1024,
    # K = 8 steps per pass, 64 x 64 regions of 8 waves x 8 rows, XCD order, the previous launch's gather inside,
    # interleaved per-cell coefficients (profiles/yee2d_tbk_batch64_kernel_and_counters.md)
    ("batch2d", "k8-8x8-xcd", 64): (2 * 423_055 + 360_374) * 1024,
}


def cpu_baseline_3d(nx, ny, nz, seconds_target=15.0):
    """The reference-order FP64 step (oracle RefSim: complex128 fields, fdtd.py:200-310 update, the dipole source
    and both port monitors, fdtd.py:411-429 order) on a slab of the 3D workload (nz planes, no z-PML), 1 core.
    The ports are sampled as planes (the reference copies whole volumes for x-normal monitors, which would make
    its step slower still)."""
    from oracle import fdtd_ref as R
    from beamz_amd.workloads import soi_mmi_3d
    mesh, d3, t, mons = soi_mmi_3d(nx, ny, nz, steps=400, z_pml=False, record_ports=True)
    ref = R.RefSim(mesh.permittivity, mesh.conductivity, mesh.dx, mesh.dy, float(t[1] - t[0]), len(t), dz=mesh.dz,
                   sources=d3.sources, monitors=mons, dtype=np.complex128, true_planes=True,
                   design_dims=(d3.width, d3.height, d3.depth))
    ref.step()                                                  # warm-up
    n, t0 = 0, time.perf_counter()
    while n < len(t) - 1:
        ref.step()
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_target:
            break
    return nx * ny * nz * n / el, n, el


def cpu_baseline_2d(sim, seconds_target=15.0):
    """The reference-order FP64 step (oracle RefSim, complex128 fields, sources + monitors) on 1 core."""
    from oracle import fdtd_ref as R
    ref = R.RefSim(sim.epsilon_r, sim.sigma, sim.dx, sim.dy, sim.dt, sim.num_steps, sources=sim.sources,
                   monitors=sim.monitors, dtype=np.complex128)
    ref.step()
    n, t0 = 0, time.perf_counter()
    while n < sim.num_steps - 1:
        ref.step()
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_target:
            break
    return sim.nx * sim.ny * n / el, n, el


def surrogate_report(a):
    """The surrogate workflow on one GPU: FDTD data generation, BF16-MFMA training (linear- and log-space
    targets), held-out error against FDTD, screening gain, and the surrogate-assisted loop against the plain
    adjoint loop (time to the plain loop's figure of merit)."""
    import torch
    from beamz_amd.surrogate import LEVELS, Surrogate
    from beamz_amd.surrogate.data import TransmissionProblem, generate, random_densities
    from beamz_amd.surrogate.optimize import (PlainAdjointLoop, SurrogateLoop, relative_error, screen, to_log,
                                              train)
    t0 = time.perf_counter()
    prob = TransmissionProblem(batch=256)
    verifier = TransmissionProblem(batch=16, seed=1)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    rho, T = generate(prob, a.samples, seed=1)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    cells = prob.side * prob.side
    gen_rate = -(-a.samples // 256) * 256 * cells * prob.steps / gen_s
    rho_v, T_v = generate(prob, 1024, seed=2)
    # FLOPs of one training sample (forward + both backward GEMMs) of the implemented network
    px = prob.px
    fwd = 0
    for lvl, (ci, co) in enumerate(LEVELS):
        hw = (px >> lvl) ** 2
        fwd += 2 * hw * 9 * (ci * co + co * co)
    res = {}
    nets = {}
    for mode in ("linear", "log"):
        net = Surrogate(seed=0)
        y = to_log(T) if mode == "log" else T
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hist = train(net, rho, y, epochs=a.epochs, batch=128, lr=5e-4)
        torch.cuda.synchronize()
        tr = time.perf_counter() - t1
        err, _ = relative_error(net, rho_v, T_v, log_space=(mode == "log"))
        nets[mode] = net
        res[mode] = {"train_s": tr, "loss_first": hist[0], "loss_last": hist[-1], "heldout_rel_l2": err,
                     "train_tflops": 3 * fwd * a.epochs * (a.samples // 128) * 128 / tr / 1e12}
    const_err = float(np.linalg.norm(T.mean(axis=0) - T_v) / np.linalg.norm(T_v))
    # a second baseline without the network: ridge regression on the density averaged over 8 x 8 pixel blocks
    def pooled(r):
        n = len(r)
        f = r.reshape(n, px // 8, 8, px // 8, 8).mean(axis=(2, 4)).reshape(n, -1)
        return np.hstack([f, np.ones((n, 1))]).astype(np.float64)
    Xp = pooled(rho)
    Wr = np.linalg.solve(Xp.T @ Xp + 10.0 * np.eye(Xp.shape[1]), Xp.T @ T)
    ridge_err = float(np.linalg.norm(pooled(rho_v) @ Wr - T_v) / np.linalg.norm(T_v))
    fom = T.mean(axis=1)
    label_stats = {"fom_mean": float(fom.mean()), "fom_cv": float(fom.std() / fom.mean()),
                   "wavelength_corr_first_last": float(np.corrcoef(T[:, 0], T[:, -1])[0, 1])}
    cands = random_densities(1024, seed=3)
    scr = {}
    for mode, net in nets.items():
        fb, fr, _ = screen(net, verifier, cands, top=16, seed=4, log_space=(mode == "log"))
        scr[mode] = {"top16_verified_fom": float(fb.mean()), "random16_verified_fom": float(fr.mean())}
    # plain adjoint loop vs surrogate-assisted loop on the broadband FOM
    plain = PlainAdjointLoop(prob, designs=16)
    t1 = time.perf_counter()
    plain_best, _ = plain.run(iterations=a.iters)
    plain_s = time.perf_counter() - t1
    best_mode = "log" if res["log"]["heldout_rel_l2"] < res["linear"]["heldout_rel_l2"] else "linear"
    loop = SurrogateLoop(nets[best_mode], verifier, log_space=(best_mode == "log"))
    order = np.argsort(-T.mean(axis=1))[:256]
    t1 = time.perf_counter()
    surr_best = loop.run(rho[order], T[order], rounds=a.rounds, target=0.99 * plain_best)
    surr_s = time.perf_counter() - t1
    reach = next((t for t, b in loop.history if b >= 0.99 * plain_best), None)
    out = {
        "metric": "Yee cell-updates/s", "value": gen_rate, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": prob.steps, "warmup": 0, "ms_per_step": gen_s / (-(-a.samples // 256) * prob.steps) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
        "data": "synthetic",
        "config": {"workload": f"surrogate: {a.samples} FDTD samples ({prob.side}x{prob.side} cells, "
                               f"{prob.px}x{prob.px} design px, {prob.steps} steps, 8 wavelengths), "
                               f"4-level BF16-MFMA CNN, {a.epochs} epochs"},
        # generation (the FDTD batch) dominates this line's wall time: algorithmic bytes at the wall-clock rate
        "roofline": {"bound": "hbm", "achieved": gen_rate * B_ALG["batch2d"] / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": gen_rate * B_ALG["batch2d"] / HBM_PEAK, "traffic": None,
                     "timing": "wall clock of the whole generation, set-up of each batch included",
                     "note_kstep": "the generation runs the K-step 2D kernel (8 steps per pass): a pass streams the "
                                   "fields and materials through HBM once for 8 steps, so the one-step algorithmic "
                                   "rate can exceed the HBM peak (frac > 1)"},
        "surrogate": {"setup_s": setup_s, "generation_s": gen_s, "train": res, "constant_predictor_rel_l2": const_err,
                      "train_roofline": {"bound": "mfma", "achieved": res[best_mode]["train_tflops"], "peak": 2500.0,
                                         "unit": "TFLOP/s", "frac": res[best_mode]["train_tflops"] / 2500.0,
                                         "flops": "forward + backward-data + backward-weight convolutions"},
                      "ridge_pooled8_rel_l2": ridge_err, "labels": label_stats,
                      "screen_1024": scr, "flop_per_sample_fwd": fwd,
                      "loop": {"plain_adjoint_best_fom": plain_best, "plain_s": plain_s, "plain_iterations": a.iters,
                               "plain_first_iteration_s": plain.history[0][0],
                               "plain_s_per_later_iteration": (plain.history[-1][0] - plain.history[0][0]) /
                               max(1, len(plain.history) - 1),
                               "plain_time_to_99pct_s": next(t for t, b in plain.history if b >= 0.99 * plain_best),
                               "surrogate_best_fom": surr_best, "surrogate_s": surr_s,
                               "surrogate_time_to_99pct_of_plain_s": reach, "surrogate_model": best_mode,
                               "one_time_cost_s": setup_s + gen_s + res[best_mode]["train_s"]}},
    }
    print(json.dumps(out))


def surrogate_train_report(a):
    """Training throughput of the surrogate's kernels: Adam steps over synthetic binary densities (batch 128),
    timed with HIP events on the network's stream after one warm-up epoch of 256 samples."""
    import torch
    from beamz_amd.surrogate import LEVELS, Surrogate
    from beamz_amd.surrogate.optimize import train
    n, epochs, px = a.samples, a.epochs, 128
    rng = np.random.default_rng(0)
    rho = (rng.random((n, px, px)) > 0.5).astype(np.float32)
    T = rng.random((n, 8)).astype(np.float32)
    net = Surrogate(seed=0)
    train(net, rho[:256], T[:256], epochs=1, batch=128)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    train(net, rho, T, epochs=epochs, batch=128)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu = e0.elapsed_time(e1) / 1e3
    fwd = sum(2 * (px >> lvl) ** 2 * 9 * (ci * co + co * co) for lvl, (ci, co) in enumerate(LEVELS))
    flops = 3 * fwd * (n // 128) * 128 * epochs
    out = {"metric": "surrogate training FLOP/s", "value": flops / wall, "unit": "FLOP/s", "n_gpus": 1,
           "steps": (n // 128) * epochs, "warmup": 2, "ms_per_step": wall / ((n // 128) * epochs) * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
           "data": "synthetic",
           "config": {"workload": f"surrogate training step: 4-level CNN {[c for _, c in LEVELS]}, 128x128 input, "
                                  f"batch 128, {n} samples x {epochs} epochs", "flop_per_sample_fwd": fwd},
           "roofline": {"bound": "mfma", "achieved": flops / gpu / 1e12, "peak": MFMA_BF16_PEAK / 1e12,
                        "unit": "TFLOP/s", "frac": flops / gpu / MFMA_BF16_PEAK, "traffic": None,
                        "timing": "HIP events around all training steps (every kernel of the step counted)"}}
    print(json.dumps(out))


class Runner:
    """What the timing loop needs from a workload: advance(k), the grid whose launches are timed."""

    def __init__(self, advance, grid, cells, desc, kernel_name, cpu):
        self.advance, self.grid, self.cells, self.desc = advance, grid, cells, desc
        self.kernel_name, self.cpu = kernel_name, cpu


def build_runner(a, total):
    import beamz_amd as bz
    from beamz_amd.workloads import batch_2d, soi_mmi_3d, topo_bend_2d, waveguide_2d
    if a.workload == "adjoint2d":
        import torch
        from beamz_amd.optimization.adjoint import AdjointProblem2D, FDTDObjective, density_to_eps
        B = a.designs
        members = [topo_bend_2d(steps=a.nt) for _ in range(B)]
        sims = [bz.FDTD(m, time=t) for m, t, _, _, _ in members]
        box, c = members[0][3], members[0][4]
        prob = AdjointProblem2D(sims, [m[2] for m in members], c, box)
        rng = np.random.default_rng(0)
        rho0 = np.clip(0.5 + 0.05 * rng.standard_normal((B, box[3], box[2])), 0.0, 1.0)
        rho = torch.tensor(rho0, dtype=torch.float32, device="cuda", requires_grad=True)
        ec, eh = 1.444 ** 2, 3.476 ** 2

        def iteration(k):
            for _ in range(int(k)):
                rho.grad = None
                J = FDTDObjective.apply(density_to_eps(rho, ec, eh, radius=3, beta=8.0), prob)
                J.sum().backward()
        # cell-updates of one gradient, sweep count set once the engines exist (2 with the design-box history of
        # every step, 3 with checkpoints and a recompute sweep)
        cells = B * sims[0].nx * sims[0].ny * a.nt
        desc = (f"adjoint gradient of a mode-overlap FOM, batch of {B} designs, {sims[0].nx}x{sims[0].ny} 2D grid, "
                f"{box[2]}x{box[3]} design px, {a.nt} steps; a step = one batched gradient")

        def cpu():
            from oracle import fdtd_ref as R
            s = sims[0]
            ref = R.RefSim(s.epsilon_r, s.sigma, s.dx, s.dy, s.dt, s.num_steps, sources=s.sources, dtype=np.complex128)
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 15.0 and n < s.num_steps:
                ref.step(); n += 1
            el = time.perf_counter() - t0
            return s.nx * s.ny * n / el, (f"{n} reference-order complex128 steps of one {s.nx}x{s.ny} design "
                                         f"({el:.1f} s); the reference's loop needs two such runs per gradient")
        r = Runner(iteration, sims[0].grid, cells, desc, None, cpu)    # kernel names: the engines exist after warmup
        r.prob = prob
        r.grids = [s.grid for s in sims]
        r.tuned = None
        r.per_step_cells = True
        return r
    if a.workload == "batch2d":
        from beamz_amd.simulation.batch import FDTDBatch
        members = batch_2d(nsims=a.nsims, steps=max(total + 4, 16))
        sims = [bz.FDTD(m, time=t) for m, t, _ in members]
        batch = FDTDBatch(sims, seg=a.seg2d, graph=a.graph, chunk=a.chunk,
                          pairs=None if a.pairs < 0 else bool(a.pairs),
                          kstep=None if a.kstep < 0 else a.kstep, krows=a.krows, kwaves=a.kwaves)
        batch.initialize()
        tuned = None
        if a.autotune:
            tuned = batch.autotune(steps=10)
        cells = batch.cells
        sides = sorted(s.nx for s in sims)
        desc = (f"{len(sims)} batched 2D TE sims, sides {sides[0]}..{sides[-1]} (seeded), 25 nm cells, "
                f"binary 2x2 um designs, 10-cell PML, 2 DFT ports x 8 wavelengths, "
                f"{'two steps per launch' if batch.engine.pairs else 'one launch per step'}")
        if batch.engine.kstep:
            K, R = batch.engine.kstep, batch.engine.krows * batch.engine.kwaves
            desc = desc.rsplit(", ", 1)[0] + f", {K} steps per pass"
            kname = f"bz_yee2d_runk (fused, batched, {K} steps per pass, 64x{R} regions, {64 - 2 * K}x{R - 2 * K} inner)"
        elif batch.engine.pairs:
            kname = f"bz_yee2d_run2 (fused, batched, two steps per launch, 60x{batch.engine.seg2} tiles)"
        else:
            kname = f"bz_yee2d_run (fused, batched, seg {batch.engine.seg}{', hipGraph' if a.graph else ''})"

        def cpu():
            big = sims[int(np.argsort([s.nx for s in sims])[len(sims) // 2])]
            v, n, el = cpu_baseline_2d(big)
            return v, f"{n} reference-order complex128 steps of the median member ({big.nx}x{big.ny}) ({el:.1f} s)"
        r = Runner(batch.advance, sims[0].grid, cells, desc, kname, cpu)
        r.grids = [s.grid for s in sims]
        r.tuned = tuned
        if batch.engine.kstep:
            r.seg_key = f"k{batch.engine.kstep}-{batch.engine.krows}x{batch.engine.kwaves}"
        else:
            r.seg_key = (f"pairs-seg{batch.engine.seg2}" if batch.engine.pairs else f"seg{batch.engine.seg}")
        r.seg_key += "-xcd" if batch.engine.xcd else ""
        r.steps_per_launch = batch.engine.kstep or (2 if batch.engine.pairs else 1)
        return r
    if a.workload == "mmi3d":
        nx, ny, nz = a.nx or 480, a.ny or 240, a.nz
        mesh, d3, t, mons = soi_mmi_3d(nx, ny, nz, steps=max(total, 16))
        sim = bz.FDTD(mesh, time=t)
        cells = nx * ny * nz
        desc = (f"3D SOI 1x2 MMI Yee step {nx}x{ny}x{nz} @25nm, 12-cell PML, dipole pulse, "
                f"2 DFT flux ports x 8 wavelengths")

        def cpu():
            nzs = max(4, nz // 8)
            v, n, el = cpu_baseline_3d(nx, ny, nzs)
            return v, (f"{n} reference-order steps (NumPy complex128 update, dipole source, both port monitors) "
                       f"of the same MMI on a {nx}x{ny}x{nzs}-cell slab ({el:.1f} s)")
    else:
        nx, ny = a.nx or 800, a.ny or 400
        d, dx, t, mon = waveguide_2d(nx, ny, steps=max(total, 16))
        sim = bz.FDTD(d, time=t, resolution=dx)
        cells = nx * ny
        desc = f"2D TM Si waveguide Yee step {nx}x{ny} @lambda/20, 20-cell PML, point pulse, 16-wavelength DFT line"

        def cpu():
            v, n, el = cpu_baseline_2d(sim)
            return v, f"{n} reference-order complex128 steps (update, source, DFT-line samples) on {nx}x{ny} ({el:.1f} s)"
    sim.grid.kernel = a.kernel
    sim.grid.kseg = a.kseg
    sim.grid.rows = a.rows
    sim.grid.tb2 = bool(a.tb2)
    sim.grid.rows2 = a.rows2
    sim.grid.kseg2 = a.kseg2
    sim.grid.waves2 = a.waves2
    sim.grid.tb3 = bool(a.tb3)
    sim.grid.rows3, sim.grid.waves3, sim.grid.kseg3 = a.rows3, a.waves3, a.kseg3
    sim.grid.seg2d = a.seg2d
    sim.grid.graph2d = a.graph
    sim.grid.kstep2d = None if a.kstep < 0 else a.kstep
    sim.grid.krows2d, sim.grid.kwaves2d = a.krows, a.kwaves
    sim.initialize_simulation(save=False, live=False)
    r = Runner(sim._advance, sim.grid, cells, desc, sim.grid.kernel_name(), cpu)
    r.steps_per_launch = None                  # 2D: set once the engine exists
    if a.workload == "mmi3d" and (a.tb2 or a.tb3):
        r.steps_per_launch = 3 if a.tb3 else 2
    r.tuned = None
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000; adjoint2d: 5 gradients)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 20; adjoint2d: 1)")
    ap.add_argument("--workload", choices=sorted(B_ALG), default="mmi3d")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=160)
    ap.add_argument("--nsims", type=int, default=64)
    ap.add_argument("--designs", type=int, default=16, help="adjoint2d: designs per batched gradient")
    ap.add_argument("--nt", type=int, default=3000, help="adjoint2d: time steps per simulation")
    ap.add_argument("--samples", type=int, default=20000, help="surrogate: FDTD training samples")
    ap.add_argument("--epochs", type=int, default=8, help="surrogate: training epochs")
    ap.add_argument("--iters", type=int, default=40, help="surrogate: plain adjoint loop iterations")
    ap.add_argument("--rounds", type=int, default=12, help="surrogate: surrogate-assisted loop rounds")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default=os.environ.get("BEAMZ_AMD_KERNEL", "fused"))
    ap.add_argument("--kseg", type=int, default=int(os.environ.get("BEAMZ_AMD_KSEG", "16")))
    ap.add_argument("--rows", type=int, default=int(os.environ.get("BEAMZ_AMD_ROWS", "2")))
    ap.add_argument("--tb2", type=int, default=1, help="3D: two steps per launch (1) or one (0)")
    ap.add_argument("--rows2", type=int, default=2, help="3D two-step kernel: y-rows per wave (2, 3, 4)")
    ap.add_argument("--kseg2", type=int, default=0, help="3D two-step kernel: z-planes per segment (0: auto)")
    ap.add_argument("--waves2", type=int, default=8, help="3D two-step kernel: waves per workgroup (8, 16)")
    ap.add_argument("--tb3", type=int, default=0, help="3D: three steps per launch (1) or not (0)")
    ap.add_argument("--rows3", type=int, default=1, help="3D three-step kernel: y-rows per wave")
    ap.add_argument("--waves3", type=int, default=16, help="3D three-step kernel: waves per workgroup")
    ap.add_argument("--kseg3", type=int, default=0, help="3D three-step kernel: z-planes per segment (0: auto)")
    ap.add_argument("--seg2d", type=int, default=None, help="rows per tile of the 2D engine (default: by size)")
    ap.add_argument("--xcd", type=int, default=1, help="2D engines: XCD-ordered tile lists (1) or row-major (0)")
    ap.add_argument("--pairs", type=int, default=-1, help="batch2d: two steps per launch (1), one (0), by size (-1)")
    ap.add_argument("--kstep", type=int, default=-1, help="2D engines: K steps per pass (0: off, -1: automatic)")
    ap.add_argument("--krows", type=int, default=None, help="2D K-step kernel: rows per wave (default: automatic)")
    ap.add_argument("--kwaves", type=int, default=None,
                    help="2D K-step kernel: waves per workgroup (default: automatic)")
    ap.add_argument("--graph", action="store_true", help="capture 2D runs as HIP graphs")
    ap.add_argument("--autotune", action="store_true",
                    help="batch2d: pick the launch shape (K steps, rows, waves; or the tile height) by timing")
    ap.add_argument("--chunk", type=int, default=96, help="batch2d: steps per engine call")
    ap.add_argument("--timing-steps", type=int, default=48,
                    help="steps timed per launch with HIP events (a multiple of the steps per launch)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (3 launches) as a counter calibration kernel")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = 5 if a.workload == "adjoint2d" else 1000
    if a.warmup is None:
        a.warmup = 1 if a.workload == "adjoint2d" else 20
    if a.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU (the engine has no multi-GPU path); run with --gpus 1"}))
        sys.exit(2)
    import torch

    if a.workload == "surrogate":
        surrogate_report(a)
        return
    if a.workload == "surrogate_train":
        surrogate_train_report(a)
        return
    total = a.warmup + a.steps + a.timing_steps
    from beamz_amd.simulation.engine import Fused2D
    Fused2D.XCD_ORDER = bool(a.xcd)
    r = build_runner(a, total)
    if a.copy_mib > 0:
        # known traffic: each launch reads and writes copy_mib MiB (FETCH_SIZE / WRITE_SIZE calibration)
        src = torch.ones(a.copy_mib * (1 << 18), device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)          # one elementwise kernel: reads and writes copy_mib MiB
        torch.cuda.synchronize()
        del src, dst
    r.advance(a.warmup)
    torch.cuda.synchronize()
    if a.workload == "adjoint2d":             # the engines exist after the warm-up gradient
        r.sweeps = 2 if r.prob.full else 3
        r.cells *= r.sweeps
        r.desc += (" (forward sweep keeping the design-box fields of every step, reverse sweep)" if r.prob.full else
                   " (forward sweep with sqrt(N) checkpoints, recompute sweep, reverse sweep)")
    if a.workload == "wg2d":                  # the 2D engine is built on the first step
        f = r.grid._f2d
        r.kernel_name = r.grid.kernel_name()
        r.steps_per_launch = f.kstep if f and f.kstep else 1
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    r.advance(a.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    value = r.cells * a.steps / wall
    # update-kernel duration: HIP events around each launch (or run of launches), on the launch stream
    r.grid.start_timing()
    r.advance(a.timing_steps)
    lt = r.grid.launch_timer()
    r.grid.timing = False
    step_kernel_s = lt if lt else gpu_s / a.steps
    b_alg = B_ALG[a.workload]
    launch_cells = r.cells if not getattr(r, "per_step_cells", False) else r.cells // (r.sweeps * a.nt)
    achieved = launch_cells * b_alg / step_kernel_s
    nx, ny = r.grid.nx, r.grid.ny
    if a.workload == "mmi3d" and a.tb3:
        key = (a.workload, f"fused-tb3-w{a.waves3}", a.rows3, r.grid.kseg3_eff(), nx, ny, r.grid.nz)
    elif a.workload == "mmi3d" and a.tb2:
        key = (a.workload, f"fused-tb2-w{a.waves2}", a.rows2, r.grid.kseg2_eff(), nx, ny, r.grid.nz)
    elif a.workload == "batch2d":
        key = (a.workload, r.seg_key, a.nsims)
    else:
        key = (a.workload, a.kernel, a.rows, a.kseg, nx, ny, r.grid.nz if a.workload == "mmi3d" else 1)
    traffic = MEASURED_TRAFFIC.get(key)
    # bytes of field and material arrays the timed kernels stream (every member's, both ping-pong sets)
    working_set = sum(g._flat.numel() * g._flat.element_size() for g in getattr(r, "grids", [r.grid]))
    if a.workload == "adjoint2d":             # the sweeps' engines exist once a gradient has run
        f = r.prob.fwd
        if f.kstep and r.prob.full:
            rc = r.prob.rec
            r.kernel_name = (f"forward sweep (design-box snapshots of every step): bz_yee2d_runk, {f.kstep} steps "
                             f"per pass, 64x{f.krows * f.kwaves} regions; reverse sweep: bz_adj2d_runk, {rc.kstep} "
                             f"steps per pass, 64x{rc.krows * rc.kwaves} regions")
        elif f.kstep:
            rc = r.prob.rec
            r.kernel_name = (f"forward sweep: bz_yee2d_runk, {f.kstep} steps per pass, 64x{f.krows * f.kwaves} "
                             f"regions; recompute sweep (with design-box snapshots): bz_yee2d_runk, and reverse "
                             f"sweep: bz_adj2d_runk, {rc.kstep} steps per pass, 64x{rc.krows * rc.kwaves} regions")
        else:
            r.kernel_name = ("forward sweep and recompute sweep (with design-box snapshots): "
                             + ("bz_yee2d_run2; reverse sweep: bz_adj2d_run2" if f.pairs
                                else "bz_yee2d_run; reverse sweep: bz_adj2d_run"))
        r.steps_per_launch = f.kstep or (2 if f.pairs else 1)
    out = {
        "metric": "Yee cell-updates/s", "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r.desc, "cells": r.cells, "kernel": r.kernel_name, "gpu_event_s": gpu_s},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel_ms": step_kernel_s * 1e3, "bytes_per_cell_update": b_alg,
                     "steps_per_launch": getattr(r, "steps_per_launch", 1), "working_set_bytes": working_set},
    }
    if working_set < INFINITY_CACHE:
        out["roofline"]["note"] = ("the working set fits in the 256 MB Infinity Cache, so the 8 TB/s HBM peak "
                                   "is not this kernel's bound; frac is the algorithmic rate over the HBM peak")
    if a.workload in ("batch2d", "wg2d") and out["roofline"]["steps_per_launch"] > 2:
        out["roofline"]["note_kstep"] = (
            "K steps per pass: a launch streams the fields and materials through HBM once for K steps, so the "
            "one-step algorithmic rate exceeds the HBM peak (frac > 1) and traffic is per launch; the kernel is "
            "bound by vector-ALU issue (profiles/yee2d_tbk_batch64_kernel_and_counters.md)")
    if r.tuned is not None:
        out["config"]["autotune_ms_per_step"] = {str(k): v for k, v in r.tuned[1].items()}
    if not a.no_cpu:
        v, sample = r.cpu()
        out["cpu_baseline"] = {"value": v, "unit": "cell-updates/s", "cores": 1, "kind": "port", "sample": sample}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
