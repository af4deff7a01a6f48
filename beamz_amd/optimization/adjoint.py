"""Exact adjoint gradients of 2D mode-overlap objectives on the GPU.

New API; the reference's topology loop (examples/4_topo.py:255-321) estimates
the gradient by the time-reversed overlap of a forward and a separate
"adjoint source" run (4_topo.py:279-286; on the GPU here as
``overlap_gradient_reference``, with the reference's adjoint source as
``reference_adjoint_source``, for comparison). Here the gradient is that of
the discrete forward map itself:

* objective: J = |a|^2 with a = sum_p c_p Ez_hat(p, w), the DFT at one
  frequency of the Ez samples a line monitor records after every step,
  projected on weights c_p (normally the conjugated, normalised mode profile
  of the output guide, ``mode_weights``);
* forward: the fused 2D engine over the whole batch of designs, with the
  full field state checkpointed every K ~ sqrt(N) steps (a multiple of 8);
  by default it advances 8 steps per launch (``bz_yee2d_runk``);
* reverse: for each segment, newest first, the segment is recomputed from its
  checkpoint by the same engine while it snapshots Ez, Hx, Hy inside the
  design box, then the reverse-time kernel walks the segment backwards,
  transposing each step and accumulating dJ/d eps_r in FP64 inside the
  design box: 8 steps per launch (``bz_adj2d_runk``) by default, else two
  (``bz_adj2d_run2``) or one (``bz_adj2d_run``). Every launch shape gives
  the same bits.

Memory: ceil(N/K) checkpoints of three field arrays plus K design-box
snapshots per design. ``FDTDObjective`` wraps this as a torch.autograd
Function of the design-box permittivity, so density filters and projections
written as torch ops chain to it.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .. import _native as N
from ..const import EPS_0, MU_0
from ..simulation.engine import Fused2D


def mode_weights(monitor, dx, dy, center, width, n_core, n_clad, wavelength):
    """c_p = conj(m_p) / |m| for the fundamental slab mode of a guide of ``width`` centred at
    ``center`` (the across-guide coordinate) sampled at the monitor's grid points, as the reference's
    topology example builds its target (examples/4_topo.py:142-164, slab_mode_source)."""
    from ..devices.mode import slab_mode_source
    pts = monitor.get_grid_points_2d(dx, dy)
    horizontal = abs(monitor.end[0] - monitor.start[0]) >= abs(monitor.end[1] - monitor.start[1])
    pos = np.array([(p[0] * dx if horizontal else p[1] * dy) for p in pts], dtype=float)
    prof, _ = slab_mode_source(x=pos - center, w=width, n_WG=n_core, n0=n_clad, wavelength=wavelength, ind_m=0, x0=0.0)
    m = np.asarray(prof, dtype=complex)
    nrm = np.linalg.norm(m)
    return np.conj(m) / (nrm if nrm > 0 else 1.0)


class AdjointProblem2D:
    """Mode-overlap objective J = sum_f s_f |a_f|^2, a_f = sum_p c_fp Ez_hat(p, w_f), and its exact
    gradient w.r.t. eps_r in a design box, for a batch of 2D simulations that share geometry of the box,
    time axis and objective monitor layout.

    sims: list of 2D ``FDTD`` objects (one additive or multiplicative source each); ``monitor`` of
    sim b is ``monitors[b]``, a line monitor with nf DFT frequencies; weights: complex c_fp, shape
    (nf, npts) or (npts,) for nf = 1 (shared); fscale: s_f (default 1); box: (x0, y0, sx, sy) design
    cells, inside the interior.
    """

    REVERSE_SEGS2 = (4, 8, 16)      # pair-tile heights tried for the recompute and reverse sweeps
    FULL_SNAPSHOT_BYTES = 48 << 30  # largest design-box history kept for every step ("auto")
    INJ_TABLE_BYTES = 256 << 20     # largest table of injection values of every step

    def __init__(self, sims, monitors, weights, box, checkpoint_every=None, seg=None, fscale=None, pairs=None,
                 kstep=None, snapshots="auto", gradient_pass=True):
        self.sims = list(sims)
        self.monitors = list(monitors)
        if len(self.monitors) != len(self.sims):
            raise ValueError("one objective monitor per simulation")
        self.c = np.atleast_2d(np.asarray(weights, dtype=np.complex128))
        self.nf = self.c.shape[0]
        for m in self.monitors:
            if m.dft_frequencies is None or len(m.dft_frequencies) != self.nf:
                raise ValueError("the objective monitor needs one DFT frequency per row of weights")
        self.fscale = np.ones(self.nf) if fscale is None else np.asarray(fscale, dtype=np.float64)
        self.box = tuple(int(v) for v in box)
        x0, y0, sx, sy = self.box
        for s in self.sims:
            if s.is_3d or s.grid.pol != "Ez":
                raise ValueError("2D simulations of the (Ez, Hx, Hy) polarisation only")
            if not (x0 >= 1 and y0 >= 1 and x0 + sx <= s.nx - 1 and y0 + sy <= s.ny - 1):
                raise ValueError("the design box must lie inside the grid interior")
        self.N = self.sims[0].num_steps
        # checkpoint interval: ceil(sqrt N), rounded up to a multiple of 8 so that the K-step launches of
        # 8 steps (the engines' automatic choice) tile every segment
        self.K = int(checkpoint_every or min(self.N, 8 * math.ceil(math.ceil(math.sqrt(self.N)) / 8)))
        self.seg = seg
        self.pairs = pairs            # two steps per launch in all three sweeps (None: by size)
        self.kstep = kstep            # K steps per launch in all three sweeps (None: automatic, 0: off)
        if snapshots not in ("auto", "full", "checkpoint"):
            raise ValueError("snapshots must be 'auto', 'full' or 'checkpoint'")
        # "full": the forward sweep keeps the design-box fields of every step and the reverse sweep reads them
        # (two sweeps per gradient); "checkpoint": fields every K steps, each segment recomputed before its
        # reverse (three sweeps, sqrt(N) memory); "auto": full while the history fits FULL_SNAPSHOT_BYTES and
        # half the free device memory
        self.snapshots = snapshots
        # gradient_pass: the reverse kernels store each design cell's lambda value per step (l2hist) and
        # bz_adj2d_grad adds the gradient terms afterwards, in the same order (the same bits); False adds them
        # inside the reverse kernels
        self.gradient_pass = bool(gradient_pass)
        self.full = None
        self._rev_tuned = False
        self.torch = self.sims[0].backend.torch
        self.dev = self.sims[0].backend.dev
        self._ready = False

    # ---------------- forward ----------------
    def _setup(self):
        T = self.torch
        grids = [s.grid for s in self.sims]
        if self._ready:
            # later evaluations reuse the engines: zero the fields and the monitor sums in place
            self._reset()
            return
        for s in self.sims:
            s.initialize_simulation(save=False, live=False)
            if s.grid_im is not None:
                raise NotImplementedError("the adjoint runs real forward fields: build the simulations with "
                                          "complex_fields=False or real sources")
        x0, y0, sx, sy = self.box
        B = len(grids)
        history = B * self.N * (3 * (sy + 1) * (sx + 1) + (sy * sx if self.gradient_pass else 0)) * 4
        if self.snapshots == "auto":
            self.full = history <= self.FULL_SNAPSHOT_BYTES and history <= T.cuda.mem_get_info(self.dev)[0] // 2
        else:
            self.full = self.snapshots == "full"
        fsnaps = None
        if self.full:
            # the design-box fields of every step, written by the forward sweep: [design][step][3][sy+1][sx+1]
            self.snap = T.zeros((B, self.N, 3, sy + 1, sx + 1), dtype=T.float32, device=self.dev)
            fsnaps = [(x0 - 1, y0 - 1, sx + 1, sy + 1, self.snap[b]) for b in range(B)]
        self.fwd = Fused2D(grids, seg=self.seg, max_steps=self.N, pairs=self.pairs, kstep=self.kstep,
                           snapshots=fsnaps)
        if self.kstep is None and self.fwd.kstep and self.N >= 64:
            # region shape of the forward sweep (when the caller left it automatic), timed once, among the shapes
            # whose K divides the checkpoint interval (every segment then runs as whole K-step launches; with the
            # full history, among all); the reverse sweep (and the recompute sweep) start from it and are timed on
            # their own in the first gradient
            shapes = [c for c in self.fwd.KCANDIDATES if self.full or self.K % c[0] == 0]
            if len(shapes) > 1:
                # 64 steps per shape: two launches of 8 steps ranked the shapes differently from run to run
                self.fwd.tune_kstep(steps=64, candidates=shapes)
                self._reset()
        elif self.fwd.pairs and self.N >= 24:
            self.fwd.tune_seg2(steps=10)
            self._reset()
        snaps = None
        if not self.full:
            self.snap = T.zeros((B, self.K, 3, sy + 1, sx + 1), dtype=T.float32, device=self.dev)
            snaps = [(x0 - 1, y0 - 1, sx + 1, sy + 1, self.snap[b]) for b in range(B)]
        # the lambda history of the design cells, slot n - snap_step0 as the snapshots
        self.l2hist = (T.empty((B, self.N if self.full else self.K, sy, sx), dtype=T.float32, device=self.dev)
                       if self.gradient_pass else None)
        # one-step tiles of the recompute engine, shared with the reverse kernel: the caller's height, else 2
        # rows (reverse sweep of 16 x 600^2 cells, 3000 steps: 107 / 110 / 128 ms at 2 / 4 / 8 rows; the
        # 8-row kernel holds 144 VGPRs and runs 3 waves per SIMD). With the full history this engine never
        # runs: it only holds the reverse sweep's tile lists.
        seg = min(self.seg, 8) if self.seg is not None else 2
        self.rec = Fused2D(grids, seg=seg, monitors=False, snapshots=snaps, pairs=self.fwd.pairs,
                           seg2=self.fwd.seg2 if self.fwd.pairs else None, kstep=self.fwd.kstep,
                           krows=self.fwd.krows, kwaves=self.fwd.kwaves)
        for g in grids:
            g._f2d = False
        self._ready = True

    def _reset(self):
        for s in self.sims:
            s.grid.reset_fields()
            s.t, s.current_step = 0, 0
        for buf in (self.fwd.series, self.fwd.dft):
            if buf is not None:
                buf.zero_()

    def set_eps(self, eps_box):
        """eps_box: (B, sy, sx) relative permittivity of the design boxes (torch, on the device)."""
        x0, y0, sx, sy = self.box
        e = eps_box.to(device=self.dev, dtype=self.torch.float32)
        for b, s in enumerate(self.sims):
            s.grid.eps[y0:y0 + sy, x0:x0 + sx] = e[b]

    def forward(self, eps_box):
        """Run every simulation with the given design permittivity; returns J (B,) FP64 torch tensor."""
        T = self.torch
        self._setup()          # first call: builds the engines; later calls: resets fields and monitor sums
        self.set_eps(eps_box)
        grids = [s.grid for s in self.sims]
        if self.full:
            # one run of every step; the engine writes the design-box history as it goes
            self.fwd.run(0, self.N, coef=True)
            for s in self.sims:
                s.current_step = self.N
            return self._objective()
        nseg = -(-self.N // self.K)
        if getattr(self, "_ckpt_buf", None) is None or len(self._ckpt_buf[0]) != nseg:
            # one buffer per grid, reused by every evaluation: (segment, component, ny, ld)
            self._ckpt_buf = [T.empty((nseg, 3) + tuple(g.f["Ez"].shape), dtype=T.float32, device=self.dev)
                              for g in grids]
        self.ckpt = [[buf[k] for buf in self._ckpt_buf] for k in range(nseg)]
        n = 0
        for k in range(nseg):
            # the segment's initial fields, all grids in one multi-tensor copy
            T._foreach_copy_([t for ck in self.ckpt[k] for t in ck.unbind(0)],
                             [g.f[c] for g in grids for c in ("Ez", "Hx", "Hy")])
            cnt = min(self.K, self.N - n)
            self.fwd.run(n, cnt, coef=(k == 0))       # the materials are fixed for the whole sweep
            n += cnt
        for s in self.sims:
            s.current_step = self.N
        return self._objective()

    def _objective(self):
        T = self.torch
        if getattr(self, "_c_dev", None) is None:
            # device copies made once: a copy from pageable host memory waits for the queued sweep to finish
            self._c_dev = T.from_numpy(self.c).to(self.dev)
            self._fs_dev = T.from_numpy(self.fscale).to(self.dev)
        c, fs = self._c_dev, self._fs_dev
        self.a = T.stack([self._overlap(s, m, c) for s, m in zip(self.sims, self.monitors)])   # (B, nf)
        self.Jf = self.a.real ** 2 + self.a.imag ** 2
        self.J = (self.Jf * fs).sum(dim=1)
        return self.J

    def _overlap(self, sim, mon, c):
        plan = next(p for p in sim._plans if p.monitor is mon)
        npts = plan.layout[1]
        z = plan.dft[:, :npts, :]                  # (nf, points, 2): Ez entries come first
        zc = self.torch.complex(z[..., 0], z[..., 1])
        return (c * zc).sum(dim=1)

    # ---------------- reverse ----------------
    def _adjoint_descs(self, lam, grad, inj):
        T = self.torch
        x0, y0, sx, sy = self.box
        descs = []
        for b, s in enumerate(self.sims):
            g = s.grid
            d = N.Adj2D()
            d.nx, d.ny, d.ld = g.nx, g.ny, g.ld
            for p in (0, 1):
                d.le[p], d.lhx[p], d.lhy[p] = (lam[b][p][q].data_ptr() for q in range(3))
            d.eps, d.sig = g.eps.data_ptr(), g.sig.data_ptr()
            d.kh = (MU_0 / EPS_0) * g.dt / (2.0 * MU_0)
            d.kbh, d.ke, d.kbe = g.dt / MU_0, g.dt / (2.0 * EPS_0), g.dt / EPS_0
            d.rdx, d.rdy = 1.0 / g.dx, 1.0 / g.dy
            src_desc = self.rec.descs[b]
            d.src_x0, d.src_y0, d.src_sx, d.src_sy = src_desc.src_x0, src_desc.src_y0, src_desc.src_sx, src_desc.src_sy
            d.src_mul = src_desc.src_mul
            d.box_x0, d.box_y0, d.box_sx, d.box_sy = x0, y0, sx, sy
            d.snap = self.snap[b].data_ptr()
            d.grad = grad[b].data_ptr()
            d.l2hist = self.l2hist[b].data_ptr() if self.l2hist is not None else None
            ix0, iy0, isx, isy, w = inj[b][:5]
            d.inj_x0, d.inj_y0, d.inj_sx, d.inj_sy = ix0, iy0, isx, isy
            d.inj_w = w.data_ptr()
            tab = inj[b][5] if len(inj[b]) > 5 else None
            d.inj_tab = tab.data_ptr() if tab is not None else None
            d.phasor = self._phasor[b].data_ptr()
            d.nf = self.nf
            # the coefficient arrays of the engine that last ran these materials: the forward engine with the full
            # history, else the recompute engine (it refreshes them at the start of every segment)
            eng = self.fwd if self.full else self.rec
            cf = eng.coefs[b] if eng.kstep and eng.coefs is not None else None
            d.coef = cf.data_ptr() if cf is not None else None
            descs.append(d)
        return descs

    def _injection_geometry(self, b):
        """The injection box of design b's objective monitor (its Ez points inside the grid) and, per box cell,
        the indices of the points on it (padded with an index of a zero entry): (ix0, iy0, isx, isy, idx, c)
        with idx (cells, points per cell) and c the weights c_fp of those points, both on the device."""
        T = self.torch
        s, m = self.sims[b], self.monitors[b]
        pts = np.array(m.get_grid_points_2d(s.dx, s.dy), dtype=np.int64)
        xs, ys = pts[:, 0], pts[:, 1]
        ok = (xs >= 0) & (xs < s.nx) & (ys >= 0) & (ys < s.ny)
        xs, ys, c = xs[ok], ys[ok], self.c[:, ok]
        ix0, iy0 = int(xs.min()), int(ys.min())
        isx, isy = int(xs.max()) - ix0 + 1, int(ys.max()) - iy0 + 1
        cell = (ys - iy0) * isx + (xs - ix0)
        per = np.bincount(cell, minlength=isy * isx)
        idx = np.full((isy * isx, max(1, int(per.max()))), len(xs), dtype=np.int64)
        fill = np.zeros(isy * isx, dtype=np.int64)
        for p, q in enumerate(cell):           # points of a cell in point order
            idx[q, fill[q]] = p
            fill[q] += 1
        return ix0, iy0, isx, isy, T.from_numpy(idx).to(self.dev), T.from_numpy(np.ascontiguousarray(c)).to(self.dev)

    def _reverse_state(self):
        """Buffers and descriptors of the reverse sweep (adjoint state, gradient, injection boxes), built on the
        first gradient and reused: a gradient then needs no host work that waits for the forward sweep."""
        if getattr(self, "_rs", None) is not None:
            return self._rs
        T = self.torch
        x0, y0, sx, sy = self.box
        grids = [s.grid for s in self.sims]
        B = len(grids)
        lam = [[T.zeros((3,) + tuple(g.box), dtype=T.float32, device=self.dev) for _ in (0, 1)] for g in grids]
        grad = T.zeros((B, sy, sx), dtype=T.float64, device=self.dev)
        geo = [self._injection_geometry(b) for b in range(B)]
        # the injection values of every step and cell, tabulated once per gradient (bz_adj2d_injtab) while they
        # fit INJ_TABLE_BYTES: the K-step reverse kernel then loads one number a step ahead instead of summing
        # weights times phasors when it needs them
        tab_bytes = sum(self.N * g[2] * g[3] * 4 for g in geo)
        tabled = self.INJ_TABLE_BYTES >= tab_bytes
        inj = [(ix0, iy0, isx, isy, T.zeros((self.nf, isy, isx, 2), dtype=T.float64, device=self.dev),
                T.zeros((self.N, isy, isx), dtype=T.float32, device=self.dev) if tabled else None)
               for ix0, iy0, isx, isy, _, _ in geo]
        self._phasor = []
        for b, s in enumerate(self.sims):
            plan = next(p for p in s._plans if p.monitor is self.monitors[b])
            if len(plan.record_steps) != self.N:
                raise ValueError("the objective monitor must record every step")
            self._phasor.append(plan.phasor.contiguous())     # (N, nf) complex as (N, 2 nf) FP64
        descs = self._adjoint_descs(lam, grad, inj)
        ddev = T.empty(len(descs) * C.sizeof(N.Adj2D), dtype=T.uint8, device=self.dev)
        self._rs = dict(lam=lam, grad=grad, geo=geo, inj=inj, descs=descs, ddev=ddev, tabled=tabled,
                        max_inj=max(i[2] * i[3] for i in inj), fs=T.from_numpy(self.fscale).to(self.dev))
        self._upload_descs()
        return self._rs

    def _upload_descs(self):
        """Stream-ordered upload of the reverse descriptors from pinned memory (the host does not wait)."""
        T = self.torch
        rs = self._rs
        raw = T.frombuffer(bytearray(b"".join(bytes(d) for d in rs["descs"])), dtype=T.uint8).pin_memory()
        rs["ddev"].copy_(raw, non_blocking=True)

    def gradient(self):
        """dJ/d eps_r over the design boxes, (B, sy, sx) FP64 torch tensor, for the last forward."""
        T = self.torch
        if not self._ready or not (self.full or hasattr(self, "ckpt")) or getattr(self, "J", None) is None:
            raise RuntimeError("call forward() first")
        x0, y0, sx, sy = self.box
        grids = [s.grid for s in self.sims]
        B = len(grids)
        rs = self._reverse_state()
        lam, grad, descs, ddev, max_inj = rs["lam"], rs["grad"], rs["descs"], rs["ddev"], rs["max_inj"]
        for pair in lam:
            for t in pair:
                t.zero_()
        grad.zero_()
        # injection weights 2 s_f conj(a_f) c_fp of this forward's objective, summed per box cell in point order,
        # on the device (no wait for the forward sweep)
        for b, ((ix0, iy0, isx, isy, idx, c), (_, _, _, _, box, _)) in enumerate(zip(rs["geo"], rs["inj"])):
            w = 2.0 * (rs["fs"] * self.a[b].conj())[:, None] * c                 # (nf, points)
            w = T.cat([w, T.zeros_like(w[:, :1])], dim=1)
            acc = w[:, idx[:, 0]]
            for k in range(1, idx.shape[1]):
                acc = acc + w[:, idx[:, k]]
            box.copy_(T.view_as_real(acc).reshape(self.nf, isy, isx, 2))
        if not self.full:
            for d in descs:
                d.snap_step0 = 0
            self._upload_descs()
        st = grids[0].stream
        if rs["tabled"]:
            N.call("bz_adj2d_injtab", ddev.data_ptr(), B, max_inj, 1, self.N, st)

        first = [True]

        def reverse(n_hi, cnt, flip):
            """Reverse steps n_hi-1 .. n_hi-cnt on the recompute engine's tile lists, then (with the gradient
            pass) their gradient terms; returns flip."""
            rec = self.rec
            if rec.kstep:
                N.call("bz_adj2d_runk", ddev.data_ptr(), B, rec.tilesk.data_ptr(), rec.ntilesk, rec.krows,
                       rec.kwaves, rec.kstep, rec.tiles.data_ptr(), rec.ntiles, rec.seg, n_hi, cnt, flip, st)
                flip = (flip + (cnt // rec.kstep) * (rec.kstep - 1)) & 1
            elif rec.pairs:
                N.call("bz_adj2d_run2", ddev.data_ptr(), B, rec.tiles2.data_ptr(), rec.ntiles2,
                       rec.seg2, rec.tiles.data_ptr(), rec.ntiles, rec.seg, n_hi, cnt, flip, st)
                flip ^= (cnt // 2) & 1
            else:
                N.call("bz_adj2d_run", ddev.data_ptr(), B, rec.tiles.data_ptr(), rec.ntiles, rec.seg, n_hi, cnt, st)
            if self.l2hist is not None:
                N.call("bz_adj2d_grad", ddev.data_ptr(), B, sx * sy, n_hi - cnt, n_hi, st)
            return flip

        def clear():
            for pair in lam:
                for t in pair:
                    t.zero_()
            grad.zero_()

        if self.full:
            # the forward sweep left the design-box fields of every step (snapshot slot n = step n): no
            # recompute, one reverse sweep over all steps
            if self.kstep is None and self.rec.kstep and not self._rev_tuned and self.N >= 192:
                ms = {}
                nt = 192 if self.N >= 384 else 96                     # a multiple of every candidate's K
                # steps from the middle of the history: the last steps' snapshots, written just before, may
                # still sit in the Infinity Cache, which favours shorter launches there
                mid = max(nt, self.N // 2)
                for shape in self.rec.KCANDIDATES:
                    self.rec.set_kstep(*shape)
                    reverse(mid, nt, 0)                               # warm-up
                    for _ in range(2):                                # the faster of two timings
                        e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
                        e0.record()
                        reverse(mid, nt, 0)
                        e1.record()
                        T.cuda.synchronize(self.dev)
                        ms[shape] = min(ms.get(shape, 1e30), e0.elapsed_time(e1))
                # within 3 % of the fastest, the most steps per launch wins: on 16 designs of 600 x 600 cells,
                # 192-step windows time K = 6 about 1.5 % faster than K = 8, while whole 3000-step sweeps with
                # K = 8 took 68.0-69.3 ms per gradient against 68.8-72.8 ms with K = 6
                best = min(ms.values())
                self.rec.set_kstep(*max((c for c in ms if ms[c] <= 1.03 * best), key=lambda c: (c[0], -ms[c])))
                self.rev_timings = ms
                clear()
            self._rev_tuned = True
            N.call("bz_adj2d_inject", ddev.data_ptr(), B, max_inj, self.N, st)
            reverse(self.N, self.N, 0)
            return grad.clone()

        def segment(k, flip):
            """Recompute segment k from its checkpoint (snapshots on), then walk it backwards; returns flip."""
            n0 = k * self.K
            cnt = min(self.K, self.N - n0)
            T._foreach_copy_([g.f[c] for g in grids for c in ("Ez", "Hx", "Hy")],
                             [t for ck in self.ckpt[k] for t in ck.unbind(0)])
            self.rec.run(n0, cnt, coef=first[0])      # coefficients once per gradient: the materials are fixed
            first[0] = False
            for d in descs:
                d.snap_step0 = n0
            self._upload_descs()
            return reverse(n0 + cnt, cnt, flip)

        nseg = len(self.ckpt)
        if self.kstep is None and self.rec.kstep and not self._rev_tuned and nseg > 1:
            # the recompute and reverse sweeps share the K-step tiles, and the reverse kernel ranks the region
            # shapes differently from the forward one: time one whole segment of both per shape whose K divides
            # the checkpoint interval and keep the fastest (the adjoint state and gradient are cleared below)
            shapes = [c for c in self.rec.KCANDIDATES if self.K % c[0] == 0]
            k_t = nseg - 1 if self.N - (nseg - 1) * self.K == self.K else nseg - 2
            if len(shapes) > 1 and k_t >= 0:
                ms = {}
                for shape in shapes:
                    self.rec.set_kstep(*shape)
                    segment(k_t, 0)                                   # warm-up
                    e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
                    e0.record()
                    segment(k_t, 0)
                    e1.record()
                    T.cuda.synchronize(self.dev)
                    ms[shape] = e0.elapsed_time(e1)
                self.rec.set_kstep(*min(ms, key=ms.get))
                self.rev_timings = ms
                clear()
            self._rev_tuned = True
        if self.rec.pairs and not self.rec.kstep and not self._rev_tuned and nseg > 1:
            # the recompute and reverse sweeps share the pair tiles: time one segment of both per tile height
            # and keep the fastest (the adjoint state and gradient are cleared again below)
            ms = {}
            for seg2 in self.REVERSE_SEGS2:
                self.rec.set_seg2(seg2)
                segment(nseg - 1, 0)                                  # warm-up
                e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
                e0.record()
                segment(nseg - 1, 0)
                e1.record()
                T.cuda.synchronize(self.dev)
                ms[seg2] = e0.elapsed_time(e1)
            self.rec.set_seg2(min(ms, key=ms.get))
            self._rev_tuned = True
            clear()
        N.call("bz_adj2d_inject", ddev.data_ptr(), B, max_inj, self.N, st)
        flip = 0                    # multi-step launches move the adjoint state across its two buffers
        for k in range(nseg - 1, -1, -1):
            flip = segment(k, flip)
        return grad.clone()


def reference_adjoint_source(design, monitor, dx, dy, target, signal, wavelength, transmission_weight=0.5,
                             mode_weight=0.5):
    """The adjoint source of the reference's topology example (examples/4_topo.py:214-233): a ModeSource line
    through the monitor's centre, as long as the monitor's line, injecting against the output direction
    (-y for a horizontal monitor, as there; -x for a vertical one) without direction enforcement, whose sample
    amplitudes are transmission_weight + mode_weight * target(s), the target mode samples interpolated along
    the line from the monitor's grid points (0 outside them). The mode solve only places the samples (its
    amplitudes are replaced), so the analytical slab solver is used."""
    from ..devices.sources import ModeSource
    pts = monitor.get_grid_points_2d(dx, dy)
    horizontal = abs(monitor.end[0] - monitor.start[0]) >= abs(monitor.end[1] - monitor.start[1])
    axis = 0 if horizontal else 1
    pos = np.array([p[axis] * (dx if horizontal else dy) for p in pts], dtype=float)
    length = float(np.hypot(monitor.end[0] - monitor.start[0], monitor.end[1] - monitor.start[1]))
    center = ((monitor.start[0] + monitor.end[0]) / 2, (monitor.start[1] + monitor.end[1]) / 2)
    src = ModeSource(design=design, position=center, width=length, wavelength=wavelength, signal=signal,
                     direction="-y" if horizontal else "-x", num_modes=1, mode_solver="analytical")
    src.enforce_direction = False
    tgt = np.asarray(target, dtype=complex)
    prof = src.mode_profiles[0]
    at = np.array([p[1 + axis] for p in prof], dtype=float)
    xpos = pos
    re = np.interp(at, xpos, tgt.real, left=0.0, right=0.0)
    im = np.interp(at, xpos, tgt.imag, left=0.0, right=0.0)
    for q, p in enumerate(prof):
        p[0] = transmission_weight + mode_weight * (re[q] + 1j * im[q])
    return src


def overlap_gradient_reference(fwd_sim, adj_sim, box, kstep=None):
    """The reference's gradient estimate (examples/4_topo.py:279-286) on the GPU:
    grad = sum_k Re(Ez_adj(k) conj(Ez_fwd(N-1-k))) over the design box, Ez(k) the field after step k of a
    forward run and of a separate run driven by the adjoint source (same time axis); with real fields the
    product needs no conjugate. No factor of dt, eps0 or the material contrast, as in the reference.

    fwd_sim, adj_sim: 2D ``FDTD`` objects of the (Ez, Hx, Hy) polarisation; box: (x0, y0, sx, sy). Each run
    records Ez inside the box after every step (the engine's snapshot box, in HBM: 12 * N * sx * sy bytes),
    and leaves its simulation's fields at step N. With complex fields (a complex source: the reference's
    complex128 path) each run also records its imaginary grid, and the product is
    Re(adj conj(fwd)) = adj_re fwd_re + adj_im fwd_im. Returns the (sy, sx) FP64 torch tensor."""
    x0, y0, sx, sy = (int(v) for v in box)
    N = fwd_sim.num_steps
    if adj_sim.num_steps != N or not np.array_equal(np.asarray(adj_sim.time), np.asarray(fwd_sim.time)):
        raise ValueError("the forward and adjoint runs must share their time axis")
    series = []
    for s in (fwd_sim, adj_sim):
        if s.is_3d or s.grid.pol != "Ez":
            raise ValueError("2D simulations of the (Ez, Hx, Hy) polarisation only")
        T = s.backend.torch
        s.initialize_simulation(save=False, live=False)
        parts = []
        for g in s._grids():
            buf = T.zeros((N, 3, sy, sx), dtype=T.float32, device=s.backend.dev)
            eng = Fused2D([g], monitors=False, snapshots=[(x0, y0, sx, sy, buf)], kstep=kstep)
            eng.run(0, N)
            g._f2d = None
            # snapshot slot q holds the Ez that step q read, i.e. the field after step q - 1
            last = g.view("Ez")[y0:y0 + sy, x0:x0 + sx]
            parts.append(T.cat([buf[1:, 0], last[None]]).double())
            del buf, eng
        s.current_step = N
        series.append(parts)
    fwd, adj = series
    out = (adj[0] * fwd[0].flip(0)).sum(dim=0)
    if len(adj) > 1 and len(fwd) > 1:
        out += (adj[1] * fwd[1].flip(0)).sum(dim=0)
    return out


class FDTDObjective:
    """torch.autograd Function: J = problem.forward(eps_box); backward = problem.gradient()."""

    @staticmethod
    def apply(eps_box, problem):
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, e):
                return problem.forward(e.detach()).to(e.dtype)

            @staticmethod
            def backward(ctx, gJ):
                g = problem.gradient().to(gJ.dtype)
                return gJ.view(-1, 1, 1) * g

        return _F.apply(eps_box)


_FILTER_TABLES = {}


def _cone_tables(radius, H, W, dev):
    """Device tables of the cone filter: the weights, the flipped weights and 1 / (K * 1) (FP64 on the host,
    rounded to FP32), cached per (radius, H, W, device)."""
    import torch
    key = (int(radius), H, W, str(dev))
    t = _FILTER_TABLES.get(key)
    if t is None:
        from numpy.lib.stride_tricks import sliding_window_view
        r = int(radius)
        yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
        K = np.maximum(0.0, float(r) - np.hypot(xx, yy))
        den = (sliding_window_view(np.pad(np.ones((H, W)), r), K.shape) * K).sum(axis=(-2, -1))
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        t = (f32(K), f32(K[::-1, ::-1]), f32(1.0 / den))
        _FILTER_TABLES[key] = t
    return t


def cone_filter_torch(rho, radius):
    """Normalised cone filter with zero padding, as ``topology.cone_filter``, on (B, H, W) FP32 device tensors,
    differentiable: the filter and its gradient run as bz_filter2d."""
    import torch

    class _Cone(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            B, H, W = x.shape
            K, _, inv = _cone_tables(radius, H, W, x.device)
            x = x.contiguous()
            y = torch.empty_like(x)
            N.call("bz_filter2d", x.data_ptr(), y.data_ptr(), B, H, W, K.data_ptr(), int(radius), None,
                   inv.data_ptr(), torch.cuda.current_stream(x.device).cuda_stream)
            return y

        @staticmethod
        def backward(ctx, g):
            B, H, W = g.shape
            _, Kf, inv = _cone_tables(radius, H, W, g.device)
            g = g.contiguous()
            gx = torch.empty_like(g)
            N.call("bz_filter2d", g.data_ptr(), gx.data_ptr(), B, H, W, Kf.data_ptr(), int(radius), inv.data_ptr(),
                   None, torch.cuda.current_stream(g.device).cuda_stream)
            return gx

    if rho.dim() != 3 or rho.dtype != torch.float32 or rho.device.type != "cuda":
        raise ValueError("cone_filter_torch takes (B, H, W) float32 tensors on the GPU")
    return _Cone.apply(rho)


def tanh_projection_torch(x, beta, eta=0.5):
    """(tanh(beta eta) + tanh(beta (x - eta))) / (tanh(beta eta) + tanh(beta (1 - eta))) (4_topo.py:107-117)."""
    import math as m
    den = m.tanh(beta * eta) + m.tanh(beta * (1.0 - eta))
    return (m.tanh(beta * eta) + (beta * (x - eta)).tanh()) / den


def density_to_eps(rho, eps_clad, eps_core, radius=3, beta=8.0, eta=0.5):
    """Design density (B, H, W) -> permittivity: cone filter, tanh projection, linear material map."""
    return eps_clad + tanh_projection_torch(cone_filter_torch(rho, radius), beta, eta) * (eps_core - eps_clad)
