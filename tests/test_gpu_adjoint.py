"""GPU: the reverse-time adjoint kernel against the FP64 adjoint oracle (itself pinned by
central differences), against FP32 finite differences on the GPU, across checkpoint
intervals, and through the torch.autograd chain of the density filters.

Tolerances: objective J within 1e-4 relative of the FP64 oracle; dJ/d eps_r within 1e-3
rel-L2 of the FP64 oracle gradient (FP32 fields and adjoint, FP64 gradient sums)."""
import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.devices.signals import gaussian_pulse
from beamz_amd.optimization.adjoint import (AdjointProblem2D, FDTDObjective, density_to_eps,
                                            overlap_gradient_reference, reference_adjoint_source)
from beamz_amd.simulation.meshing import RegularGrid
from oracle import adjoint_ref as A
from oracle import fdtd_ref as R
from oracle import topo_ref as TR

pytestmark = pytest.mark.gpu

NX, NY, DX = 70, 56, 50e-9
BOX = (25, 18, 20, 20)
F0 = bz.LIGHT_SPEED / 1.55e-6


def build(eps, sig, nsteps):
    d = bz.Design(NX * DX * (1 + 1e-9), NY * DX * (1 + 1e-9), material=bz.Material(1.0), auto_pml=False)
    dt = 0.99 * DX / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(nsteps) * dt
    st = 1 / (2 * np.pi * bz.LIGHT_SPEED * 0.3e-6 / 1.55e-6 ** 2)
    sig_t = gaussian_pulse(t, 1.0, 4 * st, st, F0, 0.0)
    src = bz.GaussianSource(position=(12 * DX, NY / 2 * DX), width=DX / 4, signal=sig_t)
    mon = bz.Monitor(design=d, start=((NX - 12) * DX, 8 * DX), end=((NX - 12) * DX, (NY - 8) * DX),
                     dft_frequencies=[F0], record_fields=False, accumulate_power=False)
    d += src
    d += mon
    mesh = RegularGrid.from_arrays(d, eps, sig, DX)
    assert mesh.permittivity.shape == (NY, NX)
    return bz.FDTD(mesh, time=t), mon, src, t


def materials(seed=0):
    rng = np.random.default_rng(seed)
    eps = np.full((NY, NX), 1.444 ** 2)
    eps[NY // 2 - 3:NY // 2 + 3, :] = 3.476 ** 2
    x0, y0, sx, sy = BOX
    eps[y0:y0 + sy, x0:x0 + sx] = 1.444 ** 2 + rng.random((sy, sx)) * (3.476 ** 2 - 1.444 ** 2)
    sig = np.zeros((NY, NX))
    for k in range(6):
        v = 4e4 * ((6 - k) / 6) ** 3
        sig[:, k] = np.maximum(sig[:, k], v); sig[:, NX - 1 - k] = np.maximum(sig[:, NX - 1 - k], v)
        sig[k, :] = np.maximum(sig[k, :], v); sig[NY - 1 - k, :] = np.maximum(sig[NY - 1 - k, :], v)
    return eps, sig, rng


def oracle_problem(sim, mon, src, eps, sig, c):
    patch = R.gaussian_patch_2d(src.position, src.width, sim.dx, sim.dy, NX, NY)
    pts = R.monitor_points_2d(mon.start, mon.end, sim.dx, sim.dy)
    return A.Problem2D(eps, sig, sim.dx, sim.dy, sim.dt, sim.num_steps, patch, src.signal, pts, c, 2 * np.pi * F0)


def test_adjoint_gradient_matches_fp64_oracle_and_checkpointing(hip):
    """The gradient against the FP64 adjoint oracle; the design-box history of every step (two sweeps) and
    checkpoints every 20 or 23 steps with a recompute sweep give the same bits."""
    torch = hip
    eps, sig, rng = materials()
    nsteps = 420
    grads = {}
    for K, mode in ((20, "checkpoint"), (23, "checkpoint"), (None, "full")):
        sim, mon, src, t = build(eps, sig, nsteps)
        npts = len(mon.get_grid_points_2d(DX, DX))
        c = np.random.default_rng(7).standard_normal(npts) + 1j * np.random.default_rng(8).standard_normal(npts)
        prob = AdjointProblem2D([sim], [mon], c, BOX, checkpoint_every=K, snapshots=mode)
        x0, y0, sx, sy = BOX
        e_box = torch.from_numpy(eps[y0:y0 + sy, x0:x0 + sx]).float().cuda()[None]
        J = prob.forward(e_box).cpu().numpy()
        grads[K] = prob.gradient().cpu().numpy()[0]
        assert prob.full == (mode == "full")
    P = oracle_problem(sim, mon, src, eps.astype(np.float32).astype(np.float64), sig.astype(np.float32).astype(np.float64), c)
    g_ref = P.gradient()
    assert abs(J[0] - P.J) <= 1e-4 * P.J, (J[0], P.J)
    gb = g_ref[y0:y0 + sy, x0:x0 + sx]
    err = np.linalg.norm(grads[20] - gb) / np.linalg.norm(gb)
    assert err < 1e-3, err
    assert np.array_equal(grads[20], grads[23])          # checkpoint interval changes nothing
    assert np.array_equal(grads[20], grads[None])        # nor keeping every step's fields


def test_adjoint_gradient_matches_gpu_finite_differences(hip):
    torch = hip
    eps, sig, rng = materials(seed=1)
    sim, mon, src, t = build(eps, sig, 360)
    npts = len(mon.get_grid_points_2d(DX, DX))
    c = np.ones(npts, dtype=complex) / np.sqrt(npts)
    prob = AdjointProblem2D([sim], [mon], c, BOX)
    x0, y0, sx, sy = BOX
    e0 = torch.from_numpy(eps[y0:y0 + sy, x0:x0 + sx]).float().cuda()[None]
    prob.forward(e0)
    g = prob.gradient().cpu().numpy()[0]
    idx = np.argsort(np.abs(g).ravel())[-4:]
    for flat in idx:
        j, i = divmod(int(flat), sx)
        h = 0.02 * float(e0[0, j, i])
        ep = e0.clone(); ep[0, j, i] += h
        em = e0.clone(); em[0, j, i] -= h
        fd = (prob.forward(ep).item() - prob.forward(em).item()) / (2 * h)
        assert abs(fd - g[j, i]) <= 0.02 * abs(g[j, i]), (j, i, fd, g[j, i])


def test_autograd_chain_through_density_filters(hip):
    torch = hip
    eps, sig, rng = materials(seed=2)
    sim, mon, src, t = build(eps, sig, 300)
    npts = len(mon.get_grid_points_2d(DX, DX))
    c = np.exp(1j * np.linspace(0, 2, npts)) / np.sqrt(npts)
    prob = AdjointProblem2D([sim], [mon], c, BOX)
    x0, y0, sx, sy = BOX
    ec, eh = 1.444 ** 2, 3.476 ** 2
    rho0 = np.clip(0.5 + 0.05 * np.random.default_rng(3).standard_normal((sy, sx)), 0, 1)
    rho = torch.tensor(rho0[None], dtype=torch.float32, device="cuda", requires_grad=True)
    e_box = density_to_eps(rho, ec, eh, radius=3, beta=8.0)
    J = FDTDObjective.apply(e_box, prob)
    J.sum().backward()
    g_rho = rho.grad.cpu().numpy()[0]
    # the same chain on the host with the oracle's filter transpose and projection derivative
    filt = TR.cone_filter(rho0, 3)
    proj, dproj = TR.tanh_projection(filt, 8.0, 0.5)
    eps_full = eps.astype(np.float32).astype(np.float64)
    eps_full[y0:y0 + sy, x0:x0 + sx] = (ec + proj * (eh - ec)).astype(np.float32)
    P = oracle_problem(sim, mon, src, eps_full, sig.astype(np.float32).astype(np.float64), c)
    g_eps = P.gradient()[y0:y0 + sy, x0:x0 + sx]
    g_ref = TR.cone_filter_T(g_eps * (eh - ec) * dproj, 3)
    err = np.linalg.norm(g_rho - g_ref) / np.linalg.norm(g_ref)
    assert err < 2e-3, err


def test_multi_frequency_adjoint_matches_fp64_oracle(hip):
    torch = hip
    eps, sig, rng = materials(seed=4)
    d = bz.Design(NX * DX * (1 + 1e-9), NY * DX * (1 + 1e-9), material=bz.Material(1.0), auto_pml=False)
    dt = 0.99 * DX / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(380) * dt
    st = 1 / (2 * np.pi * bz.LIGHT_SPEED * 0.3e-6 / 1.55e-6 ** 2)
    src = bz.GaussianSource(position=(12 * DX, NY / 2 * DX), width=DX / 4,
                            signal=gaussian_pulse(t, 1.0, 4 * st, st, F0, 0.0))
    freqs = F0 * np.array([0.96, 1.0, 1.04])
    mon = bz.Monitor(design=d, start=((NX - 12) * DX, 8 * DX), end=((NX - 12) * DX, (NY - 8) * DX),
                     dft_frequencies=freqs, record_fields=False, accumulate_power=False)
    d += src
    d += mon
    sim = bz.FDTD(RegularGrid.from_arrays(d, eps, sig, DX), time=t)
    npts = len(mon.get_grid_points_2d(DX, DX))
    c = rng.standard_normal((3, npts)) + 1j * rng.standard_normal((3, npts))
    fscale = [0.5, 1.0, 2.0]
    prob = AdjointProblem2D([sim], [mon], c, BOX, fscale=fscale)
    x0, y0, sx, sy = BOX
    J = prob.forward(torch.from_numpy(eps[y0:y0 + sy, x0:x0 + sx]).float().cuda()[None]).item()
    g = prob.gradient().cpu().numpy()[0]
    patch = R.gaussian_patch_2d(src.position, src.width, sim.dx, sim.dy, NX, NY)
    pts = R.monitor_points_2d(mon.start, mon.end, sim.dx, sim.dy)
    P = A.Problem2D(eps.astype(np.float32).astype(np.float64), sig.astype(np.float32).astype(np.float64), sim.dx,
                    sim.dy, sim.dt, sim.num_steps, patch, src.signal, pts, c, 2 * np.pi * freqs, fscale=fscale)
    gr = P.gradient()[y0:y0 + sy, x0:x0 + sx]
    assert abs(J - P.J) <= 1e-4 * P.J, (J, P.J)
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 1e-3


def test_batched_gradient_equals_single_gradients(hip):
    """Four designs in one AdjointProblem2D (one batch through all three sweeps): each design's J and
    dJ/d eps_r equal those of the design evaluated alone, bit for bit."""
    torch = hip
    eps, sig, _ = materials(seed=6)
    x0, y0, sx, sy = BOX
    rng = np.random.default_rng(10)
    boxes = np.array([eps[y0:y0 + sy, x0:x0 + sx] * (1 + 0.1 * rng.standard_normal((sy, sx))) for _ in range(4)])
    e = torch.tensor(boxes, dtype=torch.float32, device="cuda")
    built = [build(eps, sig, 330) for _ in range(4)]
    npts = len(built[0][1].get_grid_points_2d(DX, DX))
    c = np.exp(1j * np.linspace(0, 3, npts)) / np.sqrt(npts)
    prob = AdjointProblem2D([b[0] for b in built], [b[1] for b in built], c, BOX, checkpoint_every=24)
    J = prob.forward(e).cpu().numpy()
    g = prob.gradient().cpu().numpy()
    assert len(set(J.tolist())) == 4
    for b in range(4):
        s1, m1, _, _ = build(eps, sig, 330)
        p1 = AdjointProblem2D([s1], [m1], c, BOX, checkpoint_every=24)
        J1 = p1.forward(e[b:b + 1]).cpu().numpy()
        g1 = p1.gradient().cpu().numpy()
        assert np.array_equal(J1[0], J[b]) and np.array_equal(g1[0], g[b]), b


def _oracle_box_series(eps, sig, sim, sources, box):
    x0, y0, sx, sy = box
    ref = R.RefSim(eps, sig, sim.dx, sim.dy, sim.dt, sim.num_steps, sources=sources)
    out = []
    while ref.step():
        out.append(ref.F["Ez"][y0:y0 + sy, x0:x0 + sx].copy())
    return out


def test_reference_overlap_gradient_matches_oracle(hip):
    """overlap_gradient_reference: the reference's time-reversed overlap (examples/4_topo.py:279-286) of a
    forward run and a run driven by its adjoint source (4_topo.py:214-233, reference_adjoint_source), against
    the FP64 oracle's runs of the same two sources and oracle.topo_ref.overlap_gradient; 1e-4 rel-L2."""
    eps, sig, _ = materials(seed=5)
    nsteps = 260
    sim, mon, src, t = build(eps, sig, nsteps)
    sim_a, _, _, _ = build(eps, sig, nsteps)
    npts = len(mon.get_grid_points_2d(DX, DX))
    rng = np.random.default_rng(9)
    target = rng.standard_normal(npts) + 1j * rng.standard_normal(npts)
    target /= np.linalg.norm(target)
    adj = reference_adjoint_source(sim_a.design, mon, DX, DX, target, src.signal, 1.55e-6)
    sim_a.sources = [adj]
    g = overlap_gradient_reference(sim, sim_a, BOX).cpu().numpy()
    e32, s32 = eps.astype(np.float32).astype(np.float64), sig.astype(np.float32).astype(np.float64)
    fwd = _oracle_box_series(e32, s32, sim, [src], BOX)
    aser = _oracle_box_series(e32, s32, sim, [adj], BOX)
    g_ref = TR.overlap_gradient(aser, fwd)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"reference overlap gradient vs oracle: rel-L2 {err:.2e}")
    assert err < 1e-4, err


def test_c4_gradient_vs_central_differences_on_32_pixels(hip):
    """The full-size topology workload (600 x 600 cells, 120 x 120 design pixels, 3000 steps, cone filter r=3,
    tanh projection beta=8): d FOM / d rho from the adjoint + autograd chain against central differences
    (h = 1e-3 in rho) on 32 seeded pixels, the 64 perturbed forward runs advanced as one batch. FP32 fields
    limit the differences' accuracy, so the check is on the whole sampled vector: cosine > 0.99 and 2e-2
    relative L2 over the pixels."""
    torch = hip
    from beamz_amd.simulation.batch import FDTDBatch
    from beamz_amd.workloads import topo_bend_2d
    mesh, t, mon, box, c = topo_bend_2d()
    x0, y0, sx, sy = box
    ec, eh = 1.444 ** 2, 3.476 ** 2
    rng = np.random.default_rng(0)
    rho0 = np.clip(0.5 + 0.05 * rng.standard_normal((sy, sx)), 0.0, 1.0)
    sim = bz.FDTD(mesh, time=t)
    prob = AdjointProblem2D([sim], [mon], c, box)
    rho = torch.tensor(rho0[None], dtype=torch.float32, device="cuda", requires_grad=True)
    J = FDTDObjective.apply(density_to_eps(rho, ec, eh, radius=3, beta=8.0), prob)
    J.sum().backward()
    g = rho.grad.cpu().numpy()[0]
    pix = [(int(rng.integers(0, sy)), int(rng.integers(0, sx))) for _ in range(32)]
    h = 1e-3
    rhos = []
    for (j, i) in pix:
        for s in (1.0, -1.0):
            r = rho0.copy(); r[j, i] += s * h
            rhos.append(r)
    eps_all = density_to_eps(torch.tensor(np.array(rhos), dtype=torch.float32, device="cuda"), ec, eh, 3, 8.0)
    sims, mons = [], []
    for q in range(len(rhos)):
        m2, t2, mon2, _, _ = topo_bend_2d()
        s2 = bz.FDTD(m2, time=t2)
        sims.append(s2); mons.append(mon2)
    batch = FDTDBatch(sims)
    batch.initialize()
    for q, s2 in enumerate(sims):
        s2.grid.eps[y0:y0 + sy, x0:x0 + sx] = eps_all[q]
    batch.advance(batch.num_steps)
    Js = []
    for s2, m2 in zip(sims, mons):
        s2.finalize_simulation()
        a = (c * m2.dft["Ez"][0]).sum()
        Js.append(abs(a) ** 2)
    fd = np.array([(Js[2 * k] - Js[2 * k + 1]) / (2 * h) for k in range(len(pix))])
    ga = np.array([g[j, i] for (j, i) in pix])
    cos = float(fd @ ga / (np.linalg.norm(fd) * np.linalg.norm(ga)))
    err = float(np.linalg.norm(fd - ga) / np.linalg.norm(fd))
    print(f"C4 FD check: cosine {cos:.6f}, rel-L2 {err:.3e}, J {float(J):.6e}")
    assert cos > 0.99 and err < 2e-2, (cos, err)
    # the reference's own gradient estimate (examples/4_topo.py:279-286, its adjoint source of :214-233 with the
    # target mode conj(c)) at the same design, chained through the same filter and projection: compared by
    # cosine and sign, since it is a heuristic and not the derivative of the FOM
    e0 = density_to_eps(torch.tensor(rho0[None], dtype=torch.float32, device="cuda"), ec, eh, 3, 8.0)[0]
    sims2 = []
    for _ in range(2):
        m2, t2, _, _, _ = topo_bend_2d()
        s2 = bz.FDTD(m2, time=t2)
        s2.grid.eps[y0:y0 + sy, x0:x0 + sx] = e0
        sims2.append(s2)
    sim_f, sim_a = sims2
    sim_a.sources = [reference_adjoint_source(sim_a.design, mon, sim_f.dx, sim_f.dy, np.conj(c),
                                              sim_f.sources[0].signal, 1.55e-6)]
    g_ov = overlap_gradient_reference(sim_f, sim_a, box)
    rho_r = torch.tensor(rho0[None], dtype=torch.float32, device="cuda", requires_grad=True)
    (density_to_eps(rho_r, ec, eh, 3, 8.0)[0].double() * g_ov).sum().backward()
    g_ref = rho_r.grad.cpu().numpy()[0]
    gr = np.array([g_ref[j, i] for (j, i) in pix])
    cos_ref_fd = float(fd @ gr / (np.linalg.norm(fd) * np.linalg.norm(gr)))
    cos_ref_exact = float((g_ref * g).sum() / (np.linalg.norm(g_ref) * np.linalg.norm(g)))
    top = np.abs(g) >= np.quantile(np.abs(g), 0.95)
    sign_top = float(np.mean(np.sign(g_ref[top]) == np.sign(g[top])))
    sign_pix = float(np.mean(np.sign(gr) == np.sign(fd)))
    print(f"C4 reference overlap gradient: cosine vs FD {cos_ref_fd:+.4f} (sign agreement {sign_pix:.2f} on the "
          f"32 pixels), cosine vs the exact gradient {cos_ref_exact:+.4f} over the design box (sign agreement "
          f"{sign_top:.2f} on its 5 % largest entries)")
    assert np.isfinite(g_ref).all() and np.abs(g_ref).max() > 0
    assert cos > abs(cos_ref_fd)            # the exact adjoint predicts the FOM change better than the estimate


def test_pair_sweeps_give_the_same_gradient_and_engines_are_reused(hip):
    """Forward, recompute and reverse sweeps in two-step launches (checkpoint interval 23, so every segment ends
    with a single step: its snapshot slot follows the pairs', and the reverse state has moved across the two
    adjoint buffers an odd number of times) and in K-step launches (K = 5 and 8, whole launches then pairs or
    single steps) give the one-step gradient bit for bit; later evaluations on the same problem (engines
    reused, fields and monitor sums reset in place) repeat it, also after a different design was evaluated in
    between, and that in-between design's J and gradient equal those of a freshly built problem. All of it with
    checkpoints and a recompute sweep, and with the design-box history of every step kept by the forward sweep
    (one reverse call over all 300 steps)."""
    torch = hip
    eps, sig, _ = materials(seed=2)
    x0, y0, sx, sy = BOX
    e_box = torch.from_numpy(eps[y0:y0 + sy, x0:x0 + sx]).float().cuda()[None]
    out = {}
    for mode, key, pairs, kstep in [(m, *v) for m in ("checkpoint", "full") for v in
                                    (("one", False, 0), ("pairs", True, 0), ("k5", True, 5), ("k8", False, 8))]:
        sim, mon, src, t = build(eps, sig, 300)
        npts = len(mon.get_grid_points_2d(DX, DX))
        c = np.random.default_rng(7).standard_normal(npts) + 1j * np.random.default_rng(8).standard_normal(npts)
        prob = AdjointProblem2D([sim], [mon], c, BOX, checkpoint_every=23, pairs=pairs, kstep=kstep,
                                snapshots=mode)
        assert prob.snapshots == mode
        J = prob.forward(e_box).cpu().numpy()
        g = prob.gradient().cpu().numpy()
        out[mode, key] = (J, g)
        assert prob.full == (mode == "full")
        assert prob.fwd.pairs == pairs and prob.rec.pairs == pairs
        assert prob.fwd.kstep == kstep and prob.rec.kstep == kstep
        fwd, rec = prob.fwd, prob.rec
        # the in-between design on the reused engines (300 steps = 13 x 23 + 1: the newest segment is one step,
        # shorter than a K-step launch, and the coefficients must still follow the new design) against a
        # freshly built problem evaluated at that design
        Jm = prob.forward(e_box * 1.01).cpu().numpy()
        gm = prob.gradient().cpu().numpy()
        sim_f, mon_f, _, _ = build(eps, sig, 300)
        fresh = AdjointProblem2D([sim_f], [mon_f], c, BOX, checkpoint_every=23, pairs=pairs, kstep=kstep,
                                 snapshots=mode)
        Jf = fresh.forward(e_box * 1.01).cpu().numpy()
        gf = fresh.gradient().cpu().numpy()
        assert np.array_equal(Jm, Jf) and np.array_equal(gm, gf), (mode, key)
        assert not np.array_equal(gm, g), (mode, key)
        J2 = prob.forward(e_box).cpu().numpy()
        g2 = prob.gradient().cpu().numpy()
        assert prob.fwd is fwd and prob.rec is rec
        assert np.array_equal(J2, J) and np.array_equal(g2, g), (mode, key)
        out[mode, key] += (Jm, gm)
    for mode, key in out:
        for q in range(4):
            assert np.array_equal(out[mode, key][q], out["checkpoint", "one"][q]), (mode, key, q)


def test_cone_filter_kernel_and_its_gradient_match_oracle(hip):
    """bz_filter2d through cone_filter_torch: the normalised cone filter and its gradient (the transpose) against
    the oracle's FP64 cone_filter / cone_filter_T (whose adjointness the oracle tests pin), on maps that are not
    multiples of the 32 x 8 tile, several radii and a batch of maps."""
    torch = hip
    from beamz_amd.optimization.adjoint import cone_filter_torch
    rng = np.random.default_rng(9)
    for B, H, W, r in ((3, 37, 50, 3), (1, 128, 128, 3), (2, 9, 70, 1), (1, 20, 20, 5)):
        x = rng.random((B, H, W)).astype(np.float32)
        g = rng.standard_normal((B, H, W)).astype(np.float32)
        xt = torch.tensor(x, device="cuda", requires_grad=True)
        y = cone_filter_torch(xt, r)
        y.backward(torch.tensor(g, device="cuda"))
        for b in range(B):
            ref = TR.cone_filter(x[b].astype(np.float64), r)
            gref = TR.cone_filter_T(g[b].astype(np.float64), r)
            assert np.linalg.norm(y[b].detach().cpu().numpy() - ref) < 1e-6 * np.linalg.norm(ref), (B, H, W, r)
            gx = xt.grad[b].cpu().numpy()
            assert np.linalg.norm(gx - gref) < 1e-6 * np.linalg.norm(gref), (B, H, W, r)
    with pytest.raises(ValueError, match="float32"):
        cone_filter_torch(torch.zeros((1, 8, 8), dtype=torch.float64, device="cuda"), 3)


def test_gradient_pass_equals_in_kernel_accumulation(hip):
    """The reverse kernels storing each design cell's lambda value per step and bz_adj2d_grad adding the terms
    afterwards give the bits of the terms added inside the reverse kernels: with the full history and with
    checkpoints (23 steps, so segments end with single steps), for one-step, pair and K-step sweeps; the K-step
    sweeps also with the injection values tabulated per step (bz_adj2d_injtab) and without the table."""
    torch = hip
    eps, sig, _ = materials(seed=5)
    x0, y0, sx, sy = BOX
    e_box = torch.from_numpy(eps[y0:y0 + sy, x0:x0 + sx]).float().cuda()[None]
    ref = None
    for mode in ("full", "checkpoint"):
        for pairs, kstep in ((False, 0), (True, 0), (False, 8)):
            for gp, tab in ((False, True), (True, True)) + (((False, False), (True, False)) if kstep else ()):
                sim, mon, src, t = build(eps, sig, 260)
                npts = len(mon.get_grid_points_2d(DX, DX))
                c = np.exp(1j * np.linspace(0, 2, npts)) / np.sqrt(npts)
                prob = AdjointProblem2D([sim], [mon], c, BOX, checkpoint_every=23, pairs=pairs, kstep=kstep,
                                        snapshots=mode, gradient_pass=gp)
                if not tab:
                    prob.INJ_TABLE_BYTES = 0
                prob.forward(e_box)
                g = prob.gradient().cpu().numpy()
                assert (prob.l2hist is not None) == gp
                assert prob._rs["tabled"] == tab
                if ref is None:
                    ref = g
                assert np.array_equal(g, ref), (mode, pairs, kstep, gp, tab)
    assert np.abs(ref).max() > 0
