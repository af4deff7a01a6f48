"""Device state and per-step work for one FDTD simulation.

Fields live on the GPU as FP32 arrays on a padded box ([ny][ld] or
[nz][ny][ld], ld a multiple of 64 floats so every row starts on a 256-byte
boundary). Each component occupies the full box; entries outside the
reference's staggered shape are never written and stay zero, which also
provides the zero sentinel that out-of-range monitor points read
(monitors.py:287-300 returns 0.0 for them).

One step follows the reference order (beamz/simulation/fdtd.py:411-429):
H update, E update, sources, monitor sampling (+ FP64 DFT), grid power.
``Kernels`` selects the update kernels: the one-cell-per-thread pair
(bz_update_*), or the fused engine kernels when available.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ..const import EPS_0, MU_0

COMP3 = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
COMP2 = ("Ez", "Hx", "Hy")
COMP2H = ("Hz", "Ex", "Ey")     # the (Hz, Ex, Ey) polarisation, in the slots of Ez, Hx, Hy


def align(n, a=64):
    return (n + a - 1) // a * a


def ref_shapes(nx, ny, nz=None, pol="Ez"):
    """Reference field shapes (fdtd.py:151-186); the 2D (Hz, Ex, Ey) family takes those components' shapes of
    the 3D layout."""
    if nz is None:
        if pol == "Hz":
            return {"Hz": (ny - 1, nx - 1), "Ex": (ny, nx - 1), "Ey": (ny - 1, nx)}
        return {"Ez": (ny, nx), "Hx": (ny, nx - 1), "Hy": (ny - 1, nx)}
    return {"Ex": (nz, ny, nx - 1), "Ey": (nz, ny - 1, nx), "Ez": (nz - 1, ny, nx),
            "Hx": (nz - 1, ny - 1, nx), "Hy": (nz - 1, ny, nx - 1), "Hz": (nz, ny - 1, nx - 1)}


class _PointSource:
    def __init__(self, comp, off, w, mul, signal):
        self.comp, self.off, self.w, self.mul, self.signal = comp, off, w, mul, signal
        self.n = int(off.numel())


class _MonitorPlan:
    pass


class Grid:
    """Device arrays of one simulation plus the work list of one step."""

    STAGGER = 1088      # floats between consecutive arrays beyond their size (4 KiB + 256 B)

    def __init__(self, backend, eps, sigma, dx, dy, dt, dz=None, pol="Ez"):
        """pol: "Ez" (the reference's 2D family, and 3D) or "Hz" (2D (Hz, Ex, Ey) family; fused engine only)."""
        self.backend = backend
        torch = backend.torch
        self.torch = torch
        self.dev = backend.dev
        self.is_3d = eps.ndim == 3
        if pol not in ("Ez", "Hz") or (self.is_3d and pol != "Ez"):
            raise ValueError("polarization must be 'Ez', or 'Hz' for a 2D grid")
        self.pol = pol
        self.dx, self.dy, self.dz, self.dt = float(dx), float(dy), (None if dz is None else float(dz)), float(dt)
        if self.is_3d:
            self.nz, self.ny, self.nx = eps.shape
            if min(self.nx, self.ny, self.nz) < 2:
                raise ValueError("3D grids need at least 2 cells along every axis")
        else:
            self.ny, self.nx = eps.shape
            self.nz = None
            if min(self.nx, self.ny) < 2:
                raise ValueError("2D grids need at least 2 cells along every axis")
        self.ld = align(self.nx)
        box = (self.nz, self.ny, self.ld) if self.is_3d else (self.ny, self.ld)
        self.box = box
        self.names = COMP3 if self.is_3d else (COMP2H if pol == "Hz" else COMP2)
        # All arrays of one simulation (two field sets for ping-pong, permittivity, conductivity) share
        # one allocation, their bases staggered by 4 KiB + 256 B: the update kernels read every array at
        # the same element offset at once, and equally aligned bases would put those concurrent streams
        # on the same HBM channels.
        nbox = int(np.prod(box))
        stride = nbox + self.STAGGER
        narr = 2 * len(self.names) + 2
        self._flat = torch.zeros(narr * stride, dtype=torch.float32, device=self.dev)
        views = [self._flat[q * stride: q * stride + nbox].view(box) for q in range(narr)]
        nc = len(self.names)
        self._fb = [dict(zip(self.names, views[:nc])), dict(zip(self.names, views[nc:2 * nc]))]
        self._cur = 0
        self.eps = views[2 * nc]
        self.eps.fill_(1.0)
        self.sig = views[2 * nc + 1]
        self.set_materials(eps, sigma)
        self.shapes = ref_shapes(self.nx, self.ny, self.nz, pol)
        self.sources = []
        self.monitors = []
        self.power = None
        self.power_count = 0
        self.fused = None
        self.kernel = "fused"
        self._f2d = None
        self.seg2d = None       # rows per tile of the fused 2D engine (None: by grid size)
        self.graph2d = False    # capture fused 2D runs as HIP graphs
        self.kstep2d = None     # fused 2D engine: K steps per pass (0: off; None: automatic)
        self.krows2d = None     # K-step kernel: rows per wave (None: automatic)
        self.kwaves2d = None    # K-step kernel: waves per workgroup (None: automatic)
        self.kseg = 16
        self.rows = 4
        self.tb2 = True         # 3D: two steps per launch where the step list allows it
        self.kseg2 = None       # z-planes per segment of the two-step kernel (None: auto_kseg2)
        self.rows2 = 2          # y-rows per wave of the two-step kernel (2, 3 or 4)
        self.waves2 = 8         # waves per workgroup of the two-step kernel (8, or 16 with 2 rows)
        self.tb3 = False        # 3D: three steps per launch where the run allows it
        self.kseg3 = None       # z-planes per segment of the three-step kernel (None: auto_kseg3)
        self.rows3 = 1          # y-rows per wave of the three-step kernel
        self.waves3 = 16        # waves per workgroup of the three-step kernel: (rows, waves) (2, 8), (3, 8),
                                # (1, 16) or (2, 16)
        self._stage = None
        self._comb = None
        self.timing = False
        self._events = []

    @property
    def f(self):
        """Current field buffers (the fused 3D kernel ping-pongs between two sets)."""
        return self._fb[self._cur]

    def kernel_name(self):
        if self.is_3d:
            if self.kernel == "fused" and self.tb3:
                return (f"bz_yee3d_step3 (three steps per pass, {self.waves3} waves x {self.rows3} rows, "
                        f"{self.kseg3_eff()}-plane segments)")
            if self.kernel == "fused" and self.tb2:
                return (f"bz_yee3d_step2 (two steps per pass, {self.waves2} waves x {self.rows2} rows, "
                        f"{self.kseg2_eff()}-plane segments)")
            return f"bz_yee3d_step (fused z-march, {self.rows} rows/wave)" if self.kernel == "fused" else "bz_update_3d (two-pass)"
        if self.pol == "Hz":
            return "bz_yee2d_hz_run ((Hz, Ex, Ey), fused, batched)"
        if self.kernel != "fused":
            return "bz_update_h_2d + bz_update_e_2d"
        f = self._f2d
        if f and f.kstep:
            k, r = f.kstep, f.krows * f.kwaves
            return f"bz_yee2d_runk ({k} steps per pass, 64x{r} regions, {64 - 2 * k}x{r - 2 * k} inner)"
        return "bz_yee2d_run (fused, batched)"

    def start_timing(self):
        self.timing = True
        self._events = []

    def launch_timer(self):
        """Mean device time of one update launch over the timed launches, seconds (None if not measured)."""
        if not self._events:
            return None
        self.torch.cuda.synchronize(self.dev)
        ms = sum(e[0].elapsed_time(e[1]) for e in self._events)
        launches = sum(e[2] if len(e) > 2 else 1 for e in self._events)
        return ms / launches / 1e3

    # ---------------- materials / views ----------------
    def set_materials(self, eps, sigma):
        e = np.asarray(eps, dtype=np.float32)
        s = np.asarray(sigma, dtype=np.float32)
        if self.is_3d:
            self.eps[:, :, : self.nx] = self.torch.from_numpy(np.ascontiguousarray(e)).to(self.dev)
            self.sig[:, :, : self.nx] = self.torch.from_numpy(np.ascontiguousarray(s)).to(self.dev)
        else:
            self.eps[:, : self.nx] = self.torch.from_numpy(np.ascontiguousarray(e)).to(self.dev)
            self.sig[:, : self.nx] = self.torch.from_numpy(np.ascontiguousarray(s)).to(self.dev)

    def view(self, name):
        """Device view of a component in the reference's shape."""
        s = self.shapes[name]
        t = self.f[name]
        return t[: s[0], : s[1], : s[2]] if self.is_3d else t[: s[0], : s[1]]

    def offset(self, k, j, i):
        if self.is_3d:
            return (k * self.ny + j) * self.ld + i
        return j * self.ld + i

    @property
    def stream(self):
        return self.backend.stream

    def reset_fields(self):
        for fs in self._fb:
            if fs is not None:
                for t in fs.values():
                    t.zero_()
        self._cur = 0

    # ---------------- sources ----------------
    def add_point_source(self, comp, idx, w, mul, signal):
        """idx: (m, ndim) integer cell indices; w, mul: (m,) FP64 host arrays."""
        T = self.torch
        off = np.array([self.offset(*(r if self.is_3d else (0, *r))) for r in idx], dtype=np.int64)
        if len(np.unique(off)) != len(off):
            raise ValueError("point source offsets must be unique")
        src = _PointSource(comp, T.from_numpy(off).to(self.dev), T.from_numpy(np.asarray(w, np.float32)).to(self.dev),
                           None if mul is None else T.from_numpy(np.asarray(mul, np.float32)).to(self.dev),
                           np.asarray(signal))
        src.box = self._as_box(comp, np.asarray(idx), np.asarray(w, np.float64), mul)
        src.idx, src.w_host, src.mul_host = np.asarray(idx), np.asarray(w, np.float64), mul
        self.sources.append(src)
        self._f2d = None
        return src

    def _as_box(self, comp, idx, w, mul, max_cells=1 << 22):
        """Dense (add, mul) box equivalent of a point list, for the fused kernel's source epilogue."""
        if not self.is_3d or len(idx) == 0:
            return None
        lo = idx.min(axis=0); hi = idx.max(axis=0) + 1
        sz, sy, sx = (int(v) for v in (hi - lo))
        if sz * sy * sx > max_cells:
            return None
        add = np.zeros((sz, sy, sx), np.float32)
        mu = np.ones((sz, sy, sx), np.float32)
        rel = idx - lo
        add[rel[:, 0], rel[:, 1], rel[:, 2]] = w
        if mul is not None:
            mu[rel[:, 0], rel[:, 1], rel[:, 2]] = mul
        T = self.torch
        b = N.BoxSource()
        b.comp = ("Ex", "Ey", "Ez").index(comp)
        b.z0, b.y0, b.x0 = (int(v) for v in lo)
        b.sz, b.sy, b.sx = sz, sy, sx
        dadd = T.from_numpy(add.ravel()).to(self.dev)
        dmul = None if mul is None else T.from_numpy(mu.ravel()).to(self.dev)
        b.add = dadd.data_ptr()
        b.mul = 0 if dmul is None else dmul.data_ptr()
        return (b, dadd, dmul)

    # ---------------- monitors ----------------
    def add_gather(self, comps, offs, record_steps, t_all, dft_omegas, keep_series=True):
        """Plan a monitor gather: comps/offs per sample, recorded at ``record_steps``."""
        T = self.torch
        p = _MonitorPlan()
        p.n = len(offs)
        p.comp = T.from_numpy(np.asarray(comps, dtype=np.int32)).to(self.dev)
        p.off = T.from_numpy(np.asarray(offs, dtype=np.int64)).to(self.dev)
        p.record_steps = np.asarray(record_steps, dtype=np.int64)
        p.slot = {int(s): q for q, s in enumerate(p.record_steps)}
        nrec = len(p.record_steps)
        p.series = T.zeros((max(nrec, 1), p.n), dtype=T.float32, device=self.dev) if keep_series else None
        p.nfreq = 0 if dft_omegas is None else len(dft_omegas)
        if p.nfreq:
            ts = np.asarray([t_all[s] for s in p.record_steps], dtype=np.float64)
            ph = np.exp(-1j * np.outer(ts, np.asarray(dft_omegas, dtype=np.float64)))   # (nrec, nfreq)
            p.phasor = T.from_numpy(np.ascontiguousarray(ph).view(np.float64)).to(self.dev)
            p.dft = T.zeros((p.nfreq, p.n, 2), dtype=T.float64, device=self.dev)
        else:
            p.phasor = p.dft = None
        self.monitors.append(p)
        self._f2d = None
        self._stage = None
        self._comb = None
        return p

    def enable_power(self):
        self._f2d = None
        self.power = self.torch.zeros(self.box if not self.is_3d else self.box, dtype=self.torch.float64,
                                      device=self.dev)
        self.power_count = 0

    # ---------------- one step ----------------
    def _update(self):
        if self.pol == "Hz":
            raise NotImplementedError("the (Hz, Ex, Ey) polarisation runs on the fused 2D engine only")
        st = self.stream
        if self.is_3d:
            ptr = N.ptrs6([self.f[k].data_ptr() for k in COMP3])
            N.call("bz_update_3d", ptr, self.eps.data_ptr(), self.sig.data_ptr(), self.nx, self.ny, self.nz, self.ld,
                   self.dx, self.dy, self.dz, self.dt, MU_0, EPS_0, st)
        else:
            ld = self.ld
            f = self.f
            N.call("bz_update_h_2d", f["Hx"].data_ptr(), ld, f["Hy"].data_ptr(), ld, f["Ez"].data_ptr(), ld,
                   self.sig.data_ptr(), ld, self.nx, self.ny, self.dx, self.dy, self.dt, MU_0, EPS_0, st)
            N.call("bz_update_e_2d", f["Ez"].data_ptr(), ld, f["Hx"].data_ptr(), ld, f["Hy"].data_ptr(), ld,
                   self.sig.data_ptr(), ld, self.eps.data_ptr(), ld, self.nx, self.ny, self.dx, self.dy, self.dt,
                   EPS_0, st)

    def _sources(self, n):
        st = self.stream
        for s in self.sources:
            amp = float(np.real(s.signal[n]))
            f = self.f[s.comp]
            if s.mul is None:
                N.call("bz_add_points", f.data_ptr(), s.off.data_ptr(), s.w.data_ptr(), s.n, amp, st)
            else:
                N.call("bz_affine_points", f.data_ptr(), s.off.data_ptr(), s.mul.data_ptr(), s.w.data_ptr(), s.n,
                       amp, st)

    def _tables(self):
        """Gather tables: the monitor plans, or (3D) one merged table when every plan shares its record
        steps and DFT phasors, so a step costs one gather launch. Merging moves the plans' series and DFT
        sums (with what they hold) into shared buffers and leaves the plans views of them."""
        if self._comb is None:
            self._comb = False
            T = self.torch
            plans = [p for p in self.monitors if p.n > 0]
            if self.is_3d and len(plans) > 1:
                p0 = plans[0]
                same = all(np.array_equal(p.record_steps, p0.record_steps) and p.nfreq == p0.nfreq and
                           (p.series is None) == (p0.series is None) for p in plans)
                if same and p0.nfreq:
                    same = all(T.equal(p.phasor, p0.phasor) for p in plans)
                if same:
                    c = _MonitorPlan()
                    c.n = sum(p.n for p in plans)
                    c.comp = T.cat([p.comp for p in plans])
                    c.off = T.cat([p.off for p in plans])
                    c.slot, c.record_steps, c.nfreq, c.phasor = p0.slot, p0.record_steps, p0.nfreq, p0.phasor
                    c.series = None if p0.series is None else T.cat([p.series for p in plans], dim=1)
                    c.dft = T.cat([p.dft for p in plans], dim=1) if p0.nfreq else None
                    c0 = 0
                    for p in plans:
                        if c.series is not None:
                            p.series = c.series[:, c0:c0 + p.n]
                        if c.dft is not None:
                            p.dft = c.dft[:, c0:c0 + p.n, :]
                        c0 += p.n
                    c.plans = plans
                    self._comb = c
        return [self._comb] if self._comb else [p for p in self.monitors if p.n > 0]

    def _gather(self, base, t, n, off=None):
        q = t.slot.get(n)
        if q is None:
            return
        ser = 0 if t.series is None else t.series[q].data_ptr()
        o = (t.off if off is None else off).data_ptr()
        if t.nfreq:
            N.call("bz_gather", base, t.comp.data_ptr(), o, t.n, ser, t.nfreq, t.phasor[q].data_ptr(),
                   t.dft.data_ptr(), self.stream)
        else:
            N.call("bz_gather", base, t.comp.data_ptr(), o, t.n, ser, 0, None, None, self.stream)

    def _gathers(self, n):
        base = N.ptrs6([self.f[k].data_ptr() for k in self.names] + [0] * (6 - len(self.names)))
        for t in self._tables():
            self._gather(base, t, n)

    def _power(self):
        if self.power is None:
            return
        st = self.stream
        if self.is_3d:
            ptr = N.ptrs6([self.f[k].data_ptr() for k in COMP3])
            N.call("bz_grid_power_3d", self.power.data_ptr(), ptr, self.nx, self.ny, self.nz, self.ld, st)
        else:
            f = self.f
            N.call("bz_grid_power_2d", self.power.data_ptr(), f["Ez"].data_ptr(), f["Hx"].data_ptr(),
                   f["Hy"].data_ptr(), self.nx, self.ny, self.ld, st)
        self.power_count += 1

    def _fused3d(self, n):
        """One fused step; returns True if the sources were applied inside the kernel."""
        boxed = len(self.sources) <= N.MAX_BOX and all(s.box is not None for s in self.sources)
        if not boxed:
            self._slow_path("3D sources are added by separate point-list launches after each step")
        arr = (N.BoxSource * N.MAX_BOX)()
        nb = 0
        if boxed:
            for s in self.sources:
                b = s.box[0]
                b.amp = float(np.real(s.signal[n]))
                arr[nb] = b
                nb += 1
        src, dst = self._fb[self._cur], self._fb[1 - self._cur]
        ps = N.ptrs6([src[k].data_ptr() for k in COMP3])
        pd = N.ptrs6([dst[k].data_ptr() for k in COMP3])
        st = self.stream
        if self.timing:
            e0, e1 = self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)
            e0.record()
        N.call("bz_yee3d_step", ps, pd, self.eps.data_ptr(), self.sig.data_ptr(), self.nx, self.ny, self.nz, self.ld,
               self.dx, self.dy, self.dz, self.dt, MU_0, EPS_0, arr, nb, self.kseg, self.rows, st)
        if self.timing:
            e1.record()
            self._events.append((e0, e1))
        self._cur = 1 - self._cur
        return boxed

    def step(self, n):
        """Advance one step (step index n) in the reference order."""
        if not self.is_3d and self.kernel == "fused" and self._fused2d(n, 1):
            return
        if self.is_3d and self.kernel == "fused":
            if not self._fused3d(n):
                self._sources(n)
        else:
            self._update()
            self._sources(n)
        self._gathers(n)
        self._power()

    def run(self, n0, count):
        if not self.is_3d and self.kernel == "fused" and self._fused2d(n0, count):
            return
        n, end = n0, n0 + count
        while n < end:
            if end - n >= 3 and self.tb3 and self._pair_ok():
                self._fused3d_triple(n)
                n += 3
            elif end - n >= 2 and self.tb2 and self._pair_ok():
                self._fused3d_pair(n)
                n += 2
            else:
                self.step(n)
                n += 1

    # ---------------- two-step 3D path ----------------
    # workgroups of the two-step kernel one CU holds at once, by (rows, waves): VGPRs 112 -> 4 waves per
    # SIMD (two 8-wave groups); 151 and 202 VGPRs, or 16 waves -> one group
    _RESIDENT = {(2, 8): 2, (3, 8): 1, (4, 8): 1, (2, 16): 1}

    def kseg2_eff(self):
        return int(self.kseg2) if self.kseg2 else self.auto_kseg2()

    def auto_kseg2(self):
        """Segment length that splits z into as many equal segments as keep every workgroup resident at once
        (one round over the CUs, no tail): measured 0.147 ms/step at 480x240x160 (3 segments of 54) against
        0.171 ms with 32-plane segments (5 segments, 1.6 rounds)."""
        rows, waves = self.rows2, self.waves2
        gx = -(-self.nx // 60)
        gy = -(-self.ny // (waves * rows - 4))
        cus = self.torch.cuda.get_device_properties(self.dev).multi_processor_count
        capacity = cus * self._RESIDENT.get((rows, waves), 1)
        segs = max(1, capacity // (gx * gy))
        return max(2, -(-self.nz // segs))

    def kseg3_eff(self):
        return int(self.kseg3) if self.kseg3 else self.auto_kseg3()

    def auto_kseg3(self):
        """Segment length of the three-step kernel by the rule of auto_kseg2 (every shape of it holds one
        workgroup per CU)."""
        rows, waves = self.rows3, self.waves3
        gx = -(-self.nx // 58)
        gy = -(-self.ny // (waves * rows - 6))
        cus = self.torch.cuda.get_device_properties(self.dev).multi_processor_count
        segs = max(1, cus // (gx * gy))
        return max(2, -(-self.nz // segs))

    def _pair_ok(self):
        """The two- and three-step kernels apply: 3D fused path, box sources only, no grid power (needs every
        step)."""
        if not (self.is_3d and self.kernel == "fused" and (self.tb2 or self.tb3)):
            return False
        why = None
        if self.power is not None:
            why = "grid power is accumulated every step"
        elif len(self.sources) > N.MAX_BOX:
            why = f"{len(self.sources)} sources (the two-step kernel takes {N.MAX_BOX})"
        elif any(s.box is None for s in self.sources):
            why = "a source spans more than 4 M cells"
        else:
            if self._stage is None:
                self._stage = self._setup_stage()
            if self._stage is False:
                why = "the monitors span more than 8 M cells"
        if why is not None:
            self._slow_path(f"3D run takes one step per launch: {why}")
            return False
        return True

    def _slow_path(self, message):
        """Say once per grid and reason that a run left the fastest kernel."""
        seen = self.__dict__.setdefault("_slow_reported", set())
        if message not in seen:
            seen.add(message)
            from ..helpers import display_status
            display_status(message, "warning")

    def _setup_stage(self, max_cells=1 << 23):
        """Staging box covering every monitor sample, for the samples of the intermediate step."""
        T = self.torch
        plans = self._tables()
        if not plans:
            return {"plans": []}
        off = np.concatenate([p.off.cpu().numpy() for p in plans])
        plane = self.ny * self.ld
        k, rem = off // plane, off % plane
        j, i = rem // self.ld, rem % self.ld
        x0, y0, z0 = int(i.min()), int(j.min()), int(k.min())
        sx, sy, sz = int(i.max()) - x0 + 1, int(j.max()) - y0 + 1, int(k.max()) - z0 + 1
        cells = sx * sy * sz
        if cells > max_cells:
            return False
        buf = T.zeros(6 * cells, dtype=T.float32, device=self.dev)
        for p in plans:
            o = p.off.cpu().numpy()
            pk, pr = o // plane, o % plane
            pj, pi = pr // self.ld, pr % self.ld
            p.stage_off = T.from_numpy(((pk - z0) * sy + (pj - y0)) * sx + (pi - x0)).to(self.dev)
        return {"plans": plans, "buf": buf, "box": (C.c_int * 6)(x0, y0, z0, sx, sy, sz),
                "base": N.ptrs6([buf.data_ptr() + 4 * q * cells for q in range(6)])}

    def _fused3d_pair(self, n):
        """Steps n and n+1 in one launch of the two-step kernel, then the monitor samples of both."""
        arr = (N.BoxSource * N.MAX_BOX)()
        amp1 = (C.c_float * N.MAX_BOX)()
        for q, s in enumerate(self.sources):
            b = s.box[0]
            b.amp = float(np.real(s.signal[n]))
            arr[q] = b
            amp1[q] = float(np.real(s.signal[n + 1]))
        stg = self._stage
        need = any(p.slot.get(n) is not None for p in stg["plans"])
        src, dst = self._fb[self._cur], self._fb[1 - self._cur]
        ps = N.ptrs6([src[k].data_ptr() for k in COMP3])
        pd = N.ptrs6([dst[k].data_ptr() for k in COMP3])
        st = self.stream
        if self.timing:
            e0, e1 = self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)
            e0.record()
        N.call("bz_yee3d_step2", ps, pd, self.eps.data_ptr(), self.sig.data_ptr(), self.nx, self.ny, self.nz,
               self.ld, self.dx, self.dy, self.dz, self.dt, MU_0, EPS_0, arr, amp1, len(self.sources),
               stg["buf"].data_ptr() if need else None, stg["box"] if need else None, self.kseg2_eff(), self.rows2,
               self.waves2, st)
        if self.timing:
            e1.record()
            self._events.append((e0, e1, 2))
        self._cur = 1 - self._cur
        # the samples of both steps, one launch per table: step n from the staging box, step n+1 from the fields
        base1 = N.ptrs6([self.f[k].data_ptr() for k in self.names] + [0] * (6 - len(self.names)))
        for t in stg["plans"]:
            q0, q1 = t.slot.get(n), t.slot.get(n + 1)
            if q0 is None and q1 is None:
                continue
            ser = lambda q: t.series[q].data_ptr() if (t.series is not None and q is not None) else None
            ph = lambda q: t.phasor[q].data_ptr() if (t.nfreq and q is not None) else None
            N.call("bz_gather2", stg["base"], t.stage_off.data_ptr(), base1, t.comp.data_ptr(), t.off.data_ptr(),
                   t.n, int(q0 is not None), int(q1 is not None), ser(q0), ser(q1), t.nfreq, ph(q0), ph(q1),
                   t.dft.data_ptr() if t.nfreq else None, self.stream)

    def _fused3d_triple(self, n):
        """Steps n, n+1 and n+2 in one launch of the three-step kernel, then the monitor samples of all three."""
        arr = (N.BoxSource * N.MAX_BOX)()
        amp1 = (C.c_float * N.MAX_BOX)()
        amp2 = (C.c_float * N.MAX_BOX)()
        for q, s in enumerate(self.sources):
            b = s.box[0]
            b.amp = float(np.real(s.signal[n]))
            arr[q] = b
            amp1[q] = float(np.real(s.signal[n + 1]))
            amp2[q] = float(np.real(s.signal[n + 2]))
        stg = self._stage
        if stg["plans"] and "buf2" not in stg:
            stg["buf2"] = self.torch.zeros_like(stg["buf"])
            cells = stg["buf"].numel() // 6
            stg["base2"] = N.ptrs6([stg["buf2"].data_ptr() + 4 * q * cells for q in range(6)])
        need1 = any(p.slot.get(n) is not None for p in stg["plans"])
        need2 = any(p.slot.get(n + 1) is not None for p in stg["plans"])
        src, dst = self._fb[self._cur], self._fb[1 - self._cur]
        ps = N.ptrs6([src[k].data_ptr() for k in COMP3])
        pd = N.ptrs6([dst[k].data_ptr() for k in COMP3])
        st = self.stream
        if self.timing:
            e0, e1 = self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)
            e0.record()
        N.call("bz_yee3d_step3", ps, pd, self.eps.data_ptr(), self.sig.data_ptr(), self.nx, self.ny, self.nz,
               self.ld, self.dx, self.dy, self.dz, self.dt, MU_0, EPS_0, arr, amp1, amp2, len(self.sources),
               stg["buf"].data_ptr() if need1 else None, stg["buf2"].data_ptr() if need2 else None,
               stg["box"] if (need1 or need2) else None, self.kseg3_eff(), self.rows3, self.waves3, st)
        if self.timing:
            e1.record()
            self._events.append((e0, e1, 3))
        self._cur = 1 - self._cur
        # the samples of the three steps, one launch per table: steps n and n+1 from the staging boxes,
        # step n+2 from the fields
        base2 = N.ptrs6([self.f[k].data_ptr() for k in self.names] + [0] * (6 - len(self.names)))
        for t in stg["plans"]:
            q = [t.slot.get(n + s) for s in range(3)]
            if all(x is None for x in q):
                continue
            ser = lambda x: t.series[x].data_ptr() if (t.series is not None and x is not None) else None
            ph = lambda x: t.phasor[x].data_ptr() if (t.nfreq and x is not None) else None
            N.call("bz_gather3", stg["base"], stg["base2"], t.stage_off.data_ptr(), base2, t.comp.data_ptr(),
                   t.off.data_ptr(), t.n, *[int(x is not None) for x in q], *[ser(x) for x in q], t.nfreq,
                   *[ph(x) for x in q], t.dft.data_ptr() if t.nfreq else None, self.stream)

    # ---------------- fused, batched 2D path ----------------
    def invalidate(self):
        self._f2d = None
        self._stage = None
        self._comb = None

    def _fused2d(self, n0, count):
        if self._f2d is None:
            try:
                self._f2d = Fused2D([self], seg=self.seg2d, graph=self.graph2d, kstep=self.kstep2d,
                                    krows=self.krows2d, kwaves=self.kwaves2d)
            except NotImplementedError as e:
                if self.pol == "Hz":
                    raise
                from ..helpers import display_status
                display_status(f"2D run leaves the fused engine ({e}): one kernel per update, "
                               "source and monitor gather instead", "warning")
                self._f2d = False
        if self._f2d is False:
            return False
        self._f2d.run(n0, count)
        return True


def kstep_owner_tables(tiles, pts, dims, K, rh):
    """Monitor-point owner tables of the K-step kernel (bz_yee2d_runk).

    tiles: (ntiles, 3) = (sim, x-tile, y-tile) in units of the inner size (64 - 2K) x (rh - 2K), sim < 0 for
    padding; pts: (npts, 3) = (sim, comp, offset into [ny][ld]); dims: (nsims, 3) = (nx, ny, ld). Each point
    belongs to the tile whose inner cells hold it. Returns (start, ents): start[t] .. start[t+1]-1 index the
    rows of ents that tile t owns, sorted by component, ents[e] = (point index, (region row << 10) |
    (region column << 2) | comp)."""
    tiles = np.asarray(tiles, np.int64).reshape(-1, 3)
    pts = np.asarray(pts, np.int64).reshape(-1, 3)
    ix, iy = 64 - 2 * K, rh - 2 * K
    # a point of a simulation without tiles (one the batch no longer advances) has no owner and is not staged
    sim, comp = pts[:, 0], pts[:, 1]
    nxs, nys, lds = (dims[sim, c] for c in range(3))
    j, i = np.divmod(pts[:, 2], lds)
    if np.any(i >= nxs) or np.any(j >= nys):
        raise ValueError("a monitor point lies outside its grid")
    tx, ty = i // ix, j // iy
    index = {(int(a), int(b), int(c)): q for q, (a, b, c) in enumerate(tiles) if a >= 0}
    t = np.array([index.get((int(a), int(b), int(c)), -1) for a, b, c in zip(sim, tx, ty)], np.int64)
    code = ((j - (ty * iy - K)) << 10) | ((i - (tx * ix - K)) << 2) | comp
    own = np.nonzero(t >= 0)[0]
    order = own[np.lexsort((comp[own], t[own]))]
    ents = np.stack([order, code[order]], axis=1).astype(np.int32)
    start = np.searchsorted(t[order], np.arange(len(tiles) + 1)).astype(np.int32)
    return start, ents


def xcd_order(tiles, per_wg=4, nxcd=8):
    """Reorder a row-major tile list (rows of (sim, x-tile, y-tile)) for the XCDs of the GPU.

    Workgroup w takes list entries per_wg*w .. per_wg*w + per_wg - 1 and is dispatched to XCD w % nxcd,
    each XCD with its own L2. The reordered list gives XCD x one contiguous run of the original list, so
    horizontally and vertically neighbouring tiles, whose halo rows and columns share cache lines, run at
    about the same time on the same L2. Entries added to fill the last workgroups are (-1, 0, 0), which the
    kernels skip. The result is a permutation of the input plus padding: which tiles exist is unchanged."""
    tiles = np.asarray(tiles, np.int32).reshape(-1, 3)
    n = len(tiles)
    nw = -(-n // per_wg)
    cnt = np.array([nw // nxcd + (x < nw % nxcd) for x in range(nxcd)])
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    w = np.arange(nw)
    src = off[w % nxcd] + w // nxcd                       # a bijection of 0..nw-1
    idx = (src[:, None] * per_wg + np.arange(per_wg)).ravel()
    out = np.zeros((nw * per_wg, 3), np.int32)
    out[:, 0] = -1
    ok = idx < n
    out[ok] = tiles[idx[ok]]
    return out


class Fused2D:
    """Runner of bz_yee2d_run over one or several 2D grids advanced in lockstep.

    Builds the per-simulation descriptors (device array), the tile work queue
    (tiles of 62 columns x ``seg`` rows of every grid) and the combined monitor
    tables (all monitors of all grids in one gather list, one FP64 DFT
    accumulator, one series buffer; the monitors' own ``series`` / ``dft``
    become views into them). With ``graph=True`` a run of ``count`` steps is
    captured once as a HIP graph and replayed for later runs of the same length.
    """

    PAIR_MIN_CELLS = 1_000_000
    GRAPH_CHUNK = 64        # steps per captured graph (longer runs replay it)
    XCD_ORDER = True        # tile lists in xcd_order (False: plain row-major order)

    def __init__(self, grids, seg=8, graph=False, max_steps=None, monitors=True, snapshots=None, pairs=None,
                 seg2=None, xcd=None, kstep=None, krows=None, kwaves=None):
        """monitors=False runs the steps without sampling any monitor (a recomputation);
        snapshots: per grid None or (x0, y0, sx, sy, tensor [steps][3][sy][sx]) for the snapshot box;
        pairs: advance two steps per launch (bz_yee2d_run2) where a run allows it (None: from PAIR_MIN_CELLS
        cells on; never with graphs), with tiles of 60 columns x seg2 rows (None: by size); xcd: order the tile
        lists by xcd_order (None: XCD_ORDER); kstep: advance K steps per pass (bz_yee2d_runk, regions of 64 x
        krows*kwaves cells) for runs of at least K steps, the rest by pairs or single steps (0: off; None: on
        with auto_kstep's shape wherever it applies: no grid power)."""
        if not grids:
            raise ValueError("Fused2D needs at least one grid")
        self.snapshots = list(snapshots) if snapshots is not None else [None] * len(grids)
        g0 = grids[0]
        T = g0.torch
        self.torch, self.dev, self.grids = T, g0.dev, list(grids)
        self.xcd = self.XCD_ORDER if xcd is None else bool(xcd)
        self.pol = g0.pol
        for g in grids:
            if g.is_3d:
                raise ValueError("Fused2D runs 2D grids only")
            if g.dev != g0.dev:
                raise ValueError("all grids of a batch must be on one device")
            if g.pol != self.pol:
                raise ValueError("all grids of a batch must share one polarisation")
            if len(g.sources) > 255:
                raise NotImplementedError("the fused 2D engine takes at most 255 sources per simulation")
            if 4 * g.ny * g.ld >= 2 ** 32:
                raise ValueError("a 2D array exceeds 4 GiB (the engine uses 32-bit byte offsets)")
        plans = [(s, p) for s, g in enumerate(grids) for p in g.monitors if p.n > 0] if monitors else []
        # point groups: monitors with the same record steps and DFT phasors share one slot row and phasor block
        groups, pgroup = [], []
        for _, p in plans:
            for gi, q in enumerate(groups):
                if (np.array_equal(p.record_steps, q.record_steps) and p.nfreq == q.nfreq
                        and (not p.nfreq or T.equal(p.phasor, q.phasor))):
                    pgroup.append(gi)
                    break
            else:
                pgroup.append(len(groups))
                groups.append(p)
        self.keep = []
        self.active = None                    # None: every grid advances; else a bool per grid (set_active)
        self.signal_lens = {}                 # grid index -> samples of its source signal
        self.descs = []
        for g, sn in zip(grids, self.snapshots):
            self.descs.append(self._describe(g, sn))
        self.multi_src = any(d.src_sel for d in self.descs)   # some grid has several sources in its box
        self.desc_dev = T.zeros(len(grids) * C.sizeof(N.Sim2D), dtype=T.uint8, device=self.dev)
        self._desc_cache = {}
        self.shift = None
        self.set_seg(seg)
        self.step = T.zeros(1, dtype=T.int32, device=self.dev)
        self.npts, self.nfreq = 0, 0
        self.pts = self.slot = self.series = self.dft = self.phasor = self.grp = None
        self.nslot = self.nrow = 0
        self.groups = [q.record_steps for q in groups]
        if plans:
            ntot = sum(p.n for _, p in plans)
            sim = np.concatenate([np.full(p.n, s, np.int32) for s, p in plans])
            comp = T.cat([p.comp for _, p in plans]).cpu().numpy()
            off = T.cat([p.off for _, p in plans]).cpu().numpy()
            if off.max(initial=0) >= 2 ** 31:
                raise ValueError("monitor offsets exceed 32 bits")
            pts = np.stack([sim, comp, off.astype(np.int32)], axis=1).astype(np.int32)
            self.pts = T.from_numpy(pts.ravel()).to(self.dev)
            if len(groups) > 1:
                self.grp = T.from_numpy(np.repeat(np.asarray(pgroup, np.int32), [p.n for _, p in plans])).to(self.dev)
            last = max((int(rs.max()) + 1 for rs in self.groups if len(rs)), default=1)
            need = max(last, int(max_steps or 0), 1)
            self._make_slot(need)
            self.nrow = max(max(len(rs) for rs in self.groups), 1)
            self.nfreq = max(q.nfreq for q in groups)
            if self.nfreq:
                if len(groups) == 1:
                    self.phasor = groups[0].phasor
                else:
                    # [group][row][frequency] complex; frequencies a group lacks stay 0
                    self.phasor = T.zeros((len(groups), self.nrow, 2 * self.nfreq), dtype=T.float64,
                                          device=self.dev)
                    for gi, q in enumerate(groups):
                        if q.nfreq:
                            self.phasor[gi, :len(q.record_steps), :2 * q.nfreq] = q.phasor
            keep = any(p.series is not None for _, p in plans)
            self.series = T.zeros((self.nrow, ntot), dtype=T.float32, device=self.dev) if keep else None
            self.dft = T.zeros((self.nfreq, ntot, 2), dtype=T.float64, device=self.dev) if self.nfreq else None
            c0 = 0
            for _, p in plans:
                if p.series is not None:
                    p.series = self.series[:max(len(p.record_steps), 1), c0:c0 + p.n]
                if p.nfreq:
                    p.dft = self.dft[:p.nfreq, c0:c0 + p.n, :]
                c0 += p.n
            self.npts = ntot
        self.use_graph = bool(graph)
        self.graphs = {}
        if self.pol == "Hz":
            # the (Hz, Ex, Ey) family has the one-step kernel only
            if any(sn is not None for sn in self.snapshots):
                raise ValueError("snapshot boxes need the (Ez, Hx, Hy) polarisation")
            pairs, kstep = False, 0
        if pairs is None:
            # small batches are launch-latency bound, and a pair's serial row march is slower there
            # (800 x 400: 13.4 us/step in pairs against 10.2 with one step per launch)
            pairs = sum(g.nx * g.ny for g in grids) >= self.PAIR_MIN_CELLS
        self.pairs = bool(pairs) and not self.use_graph
        self.spts = None
        if self.pairs:
            if self.npts:
                self._setup_staging(pts)
            self.set_seg2(seg2)
        self._pts_host = pts if plans else None
        self.kstep, self.krows, self.kwaves = 0, 8, 8
        self.tilesk = self.ent_start = self.ents = self.stagek = None
        self.coefs = None
        self._coef_done = False
        if kstep is None and self.kstep_applies():
            kstep, r, w = self.auto_kstep()
            krows, kwaves = krows or r, kwaves or w
        if kstep and not self.kstep_applies():
            kstep = 0           # grid power needs every step: pairs or single steps instead
        if kstep:
            self.set_kstep(kstep, krows or 8, kwaves or 8)

    def _setup_staging(self, pts):
        """Per simulation, a staging box around its monitor points for the first step of each pair."""
        T = self.torch
        spts = np.full(len(pts), -1, np.int32)
        for s, g in enumerate(self.grids):
            sel = np.nonzero(pts[:, 0] == s)[0]
            if len(sel) == 0:
                continue
            j, i = np.divmod(pts[sel, 2].astype(np.int64), g.ld)
            real = pts[sel, 2] != g.offset(0, 0, g.nx - 1)       # the zero sentinel of out-of-range points
            if not real.any():
                continue
            x0, x1 = int(i[real].min()), int(i[real].max()) + 1
            y0, y1 = int(j[real].min()), int(j[real].max()) + 1
            sx, sy = x1 - x0, y1 - y0
            buf = T.zeros(2 * 3 * sx * sy, dtype=T.float32, device=self.dev)     # two slots: pair q uses q & 1
            self.keep.append(buf)
            d = self.descs[s]
            d.stg, d.stg_x0, d.stg_y0, d.stg_sx, d.stg_sy = buf.data_ptr(), x0, y0, sx, sy
            inside = real & (i >= x0) & (i < x1) & (j >= y0) & (j < y1)
            spts[sel[inside]] = ((j[inside] - y0) * sx + (i[inside] - x0)).astype(np.int32)
        self.spts = T.from_numpy(spts).to(self.dev)
        self.shift = None            # descriptors changed: upload again
        self._desc_cache = {}

    SEGS2 = (4, 8, 16, 32, 64)

    def auto_seg2(self, min_waves=49152):
        """Tallest pair tile (fewest prologue rows) that still gives min_waves tiles, else 4.

        One tile is one wave and about 8192 waves are resident at once, so 49152 tiles are six rounds:
        fewer rounds leave the last one ragged. On 64 members of 256..1024 cells a side (30 M cells) the
        measured step times were 0.150 / 0.139 / 0.146 / 0.159 / 0.177 ms for 4 / 8 / 16 / 32 / 64 rows."""
        for seg in sorted(self.SEGS2, reverse=True):
            if sum(-(-g.nx // 60) * -(-g.ny // seg) for g in self.grids) >= min_waves:
                return seg
        return min(self.SEGS2)

    def tune_seg2(self, n0=0, steps=10):
        """Time every pair-tile height of SEGS2 on `steps` steps from step n0 (HIP events on the launch stream)
        and keep the fastest; returns (best, {seg2: ms per step}). The grids' fields and the monitor sums are
        advanced by the timing runs: callers reset them afterwards."""
        T = self.torch
        if not self.pairs:
            raise ValueError("tune_seg2 needs an engine that runs pairs")
        steps += steps % 2
        timings = {}
        kstep, self.kstep = self.kstep, 0                    # time the pair kernel alone
        try:
            for seg in self.SEGS2:
                self.set_seg2(seg)
                self.run(n0, 2)                               # warm-up
                e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
                e0.record()
                self.run(n0 + 2, steps)
                e1.record()
                T.cuda.synchronize(self.dev)
                timings[seg] = e0.elapsed_time(e1) / steps
        finally:
            self.kstep = kstep
        best = min(timings, key=timings.get)
        self.set_seg2(best)
        return best, timings

    def is_active(self, s):
        return self.active is None or bool(self.active[s])

    def _advancing(self):
        return [g for s, g in enumerate(self.grids) if self.is_active(s)]

    def set_active(self, active):
        """Advance only the grids whose entry is true (None: all), e.g. when batch members with shorter time
        axes have finished: the tile lists are rebuilt without the others, whose fields, monitor sums and
        buffer parity then stay as they are (their record schedules have no rows beyond their own steps)."""
        active = None if active is None or all(active) else [bool(a) for a in active]
        if active == self.active:
            return
        if active is not None and not any(active):
            raise ValueError("set_active needs at least one active grid")
        self.active = active
        self.set_seg(self.seg)
        if self.pairs:
            self.set_seg2(self.seg2)
        if self.kstep:
            self.set_kstep(self.kstep, self.krows, self.kwaves)
        self.graphs = {}

    def set_seg2(self, seg2):
        """Tile list of the two-step kernel: tiles of 60 columns x seg2 rows."""
        T = self.torch
        self.seg2 = self.auto_seg2() if seg2 is None else int(seg2)
        if self.seg2 < 2:
            raise ValueError("pair tiles need at least 2 rows")
        rows = []
        for s, g in enumerate(self.grids):
            if not self.is_active(s):
                continue
            gx = (g.nx + 59) // 60
            gy = (g.ny + self.seg2 - 1) // self.seg2
            t = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="xy"), -1).reshape(-1, 2)
            rows.append(np.concatenate([np.full((len(t), 1), s), t], axis=1))
        tiles = np.concatenate(rows).astype(np.int32)
        if self.xcd:
            tiles = xcd_order(tiles)
        self.tiles2 = T.from_numpy(tiles.ravel()).to(self.dev)
        self.ntiles2 = len(tiles)

    KSHAPES = ((8, 4), (8, 8), (8, 16), (4, 8), (4, 16), (16, 4))     # (rows per wave, waves per workgroup)
    # (K, rows, waves) candidates of tune_kstep; the automatic shapes were the fastest measured on 64 members
    # of 256..1024 cells a side (30 M cells: 0.051 ms/step for (8, 8, 8)) and on one 800 x 400 grid with a
    # 16-wavelength DFT line (1.43 us/step for (12, 4, 16), against 1.51 for (8, 4, 8) and 1.80 for (8, 8, 8));
    # 64 x 128 regions, (8, 8, 16), fewer recomputed halo cells but one workgroup per CU: 0.057 ms/step on the
    # 64 members against 0.043 for (8, 8, 8)
    KCANDIDATES = ((8, 8, 8), (6, 8, 8), (8, 8, 4), (8, 4, 8), (12, 4, 16))

    def kstep_applies(self):
        return all(g.power is None for g in self.grids)

    def auto_kstep(self):
        """(K, rows, waves): 64 x 64 regions of 8 waves x 8 rows from PAIR_MIN_CELLS cells on; below, 64 x 64
        regions of 16 waves x 4 rows and 12 steps per pass, which halve each wave's rows and the launches."""
        big = sum(g.nx * g.ny for g in self.grids) >= self.PAIR_MIN_CELLS
        return (8, 8, 8) if big else (12, 4, 16)

    def tune_kstep(self, n0=0, steps=16, candidates=None):
        """Time every (K, rows, waves) of `candidates` (default KCANDIDATES) on about `steps` steps from step
        n0 and keep the fastest; returns (best, {shape: ms per step}). Advances the fields and the monitor sums
        (callers reset them afterwards)."""
        T = self.torch
        if not self.kstep:
            raise ValueError("tune_kstep needs an engine that runs K steps per pass")
        timings = {}
        for shape in candidates or self.KCANDIDATES:
            K, rows, waves = shape
            self.set_kstep(K, rows, waves)
            n = max(1, steps // K) * K
            self.run(n0, K)                                   # warm-up
            e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
            e0.record()
            self.run(n0 + K, n, coef=False)
            e1.record()
            T.cuda.synchronize(self.dev)
            timings[shape] = e0.elapsed_time(e1) / n
        best = min(timings, key=timings.get)
        self.set_kstep(*best)
        return best, timings

    def set_kstep(self, K, rows=8, waves=8):
        """Tile list, monitor-owner tables and sample staging of the K-steps-per-pass kernel (K 0: off).

        A tile's region is 64 columns x rows*waves rows; its inner (64-2K) x (rows*waves-2K) cells tile the
        grid, and each monitor point is sampled by the tile whose inner cells hold it."""
        T = self.torch
        K = int(K or 0)
        if not K:
            self.kstep = 0
            return
        rows, waves = int(rows), int(waves)
        if (rows, waves) not in self.KSHAPES or not 2 * K < min(64, rows * waves):
            raise ValueError(f"K steps per pass need (rows, waves) in {self.KSHAPES} and 2K < min(64, rows*waves)")
        if any(g.power is not None for g in self.grids):
            raise ValueError("K steps per pass do not accumulate grid power")
        ix, iy = 64 - 2 * K, rows * waves - 2 * K
        parts = []
        for s, g in enumerate(self.grids):
            if not self.is_active(s):
                continue
            gx, gy = -(-g.nx // ix), -(-g.ny // iy)
            t = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="xy"), -1).reshape(-1, 2)
            parts.append(np.concatenate([np.full((len(t), 1), s), t], axis=1))
        tiles = np.concatenate(parts).astype(np.int32)
        if self.xcd:
            tiles = xcd_order(tiles, per_wg=1)
        self.kstep, self.krows, self.kwaves = K, rows, waves
        self.graphs = {}                     # captured launches hold the tables' addresses
        if self.coefs is None:
            # per-cell update coefficients, computed at the start of every run (bz_yee2d_coef): the K-step kernel
            # then loads four numbers per cell instead of evaluating three divisions. The kernels address them
            # with 32-bit byte offsets; a grid too large for that keeps coef NULL and the kernels divide.
            self.coefs = [T.empty((g.ny, g.ld, 4), dtype=T.float32, device=self.dev)
                          if 16 * g.ny * g.ld < 2 ** 32 else None for g in self.grids]
            for d, c in zip(self.descs, self.coefs):
                d.coef = c.data_ptr() if c is not None else None
            self.shift = None
            self._desc_cache = {}
        self.tilesk = T.from_numpy(tiles.ravel()).to(self.dev)
        self.ntilesk = len(tiles)
        self.ent_start = self.ents = self.stagek = None
        if self.npts:
            dims = np.array([(g.nx, g.ny, g.ld) for g in self.grids], np.int64)
            start, ents = kstep_owner_tables(tiles, self._pts_host, dims, K, rows * waves)
            self.ents = T.from_numpy(ents.ravel()).to(self.dev)
            self.ent_start = T.from_numpy(start).to(self.dev)
            self.stagek = T.zeros(2 * K * self.npts, dtype=T.float32, device=self.dev)   # two slots: q & 1

    def _describe(self, g, snap=None):
        T = self.torch
        d = N.Sim2D()
        d.nx, d.ny, d.ld = g.nx, g.ny, g.ld
        d.eps, d.sig = g.eps.data_ptr(), g.sig.data_ptr()
        d.kh = (MU_0 / EPS_0) * g.dt / (2.0 * MU_0)
        d.kbh, d.ke, d.kbe = g.dt / MU_0, g.dt / (2.0 * EPS_0), g.dt / EPS_0
        d.rdx, d.rdy = 1.0 / g.dx, 1.0 / g.dy
        d.src_sx = 0
        d.src_sel, d.src_sig_len = None, 0
        if g.sources:
            # one box holds every source of the grid; with several, each cell records whose it is (the sources
            # must touch disjoint cells: each cell then takes the one affine update it takes alone)
            srcs = g.sources
            idx = np.concatenate([s.idx for s in srcs])
            if len(srcs) > 1 and len(np.unique(idx[:, 0] * (g.nx + 1) + idx[:, 1])) != len(idx):
                raise NotImplementedError("the fused 2D engine needs the sources of a simulation on disjoint cells")
            lo = idx.min(axis=0); hi = idx.max(axis=0) + 1
            sy, sx = (int(v) for v in (hi - lo))
            add = np.zeros((sy, sx), np.float32); mul = np.ones((sy, sx), np.float32)
            sel = np.zeros((sy, sx), np.uint8)
            for q, s in enumerate(srcs):
                rel = s.idx - lo
                add[rel[:, 0], rel[:, 1]] = s.w_host
                sel[rel[:, 0], rel[:, 1]] = q
                if s.mul_host is not None:
                    mul[rel[:, 0], rel[:, 1]] = s.mul_host
            has_mul = any(s.mul_host is not None for s in srcs)
            dadd = T.from_numpy(add.ravel()).to(self.dev)
            dmul = T.from_numpy(mul.ravel()).to(self.dev) if has_mul else None
            ln = min(len(s.signal) for s in srcs)
            sig = np.stack([np.real(s.signal[:ln]).astype(np.float32) for s in srcs])
            dsig = T.from_numpy(sig.ravel()).to(self.dev)
            dsel = T.from_numpy(sel.ravel()).to(self.dev) if len(srcs) > 1 else None
            self.keep += [dadd, dmul, dsig, dsel]
            d.src_y0, d.src_x0, d.src_sy, d.src_sx = int(lo[0]), int(lo[1]), sy, sx
            d.src_add = dadd.data_ptr()
            d.src_mul = dmul.data_ptr() if dmul is not None else None
            d.signal = dsig.data_ptr()
            if dsel is not None:
                d.src_sel, d.src_sig_len = dsel.data_ptr(), ln
            self.signal_lens[len(self.descs)] = ln
        d.power = g.power.data_ptr() if g.power is not None else None
        if snap is not None:
            x0, y0, sx, sy, buf = snap
            if not (0 <= x0 and 0 <= y0 and x0 + sx <= g.nx and y0 + sy <= g.ny):
                raise ValueError("snapshot box outside the grid")
            if buf.dtype != T.float32 or buf.shape[1:] != (3, sy, sx) or not buf.is_contiguous():
                raise ValueError("snapshot buffer must be contiguous FP32 [steps][3][sy][sx]")
            d.snap_x0, d.snap_y0, d.snap_sx, d.snap_sy = x0, y0, sx, sy
            d.snap = buf.data_ptr()
            self.snap_steps = min(getattr(self, "snap_steps", 1 << 62), buf.shape[0])
        return d

    SEGS = (2, 4, 8, 16)
    TILE_X = 62

    def auto_seg(self, min_waves=8192, max_seg=8):
        """Tallest tile up to max_seg rows (fewest halo rows; 16-row tiles need 170 VGPRs, 2 waves per
        SIMD) that still gives min_waves tiles (8 waves per SIMD), else 2."""
        for seg in sorted((s for s in self.SEGS if s <= max_seg), reverse=True):
            if sum(-(-g.nx // self.TILE_X) * -(-g.ny // seg) for g in self.grids) >= min_waves:
                return seg
        return min(self.SEGS)

    def set_seg(self, seg):
        """Rebuild the tile work queue with tiles of 62 columns x seg rows (seg None: auto_seg)."""
        T = self.torch
        seg = self.auto_seg() if seg is None else int(seg)
        if seg not in self.SEGS:
            raise ValueError(f"tile height must be one of {self.SEGS}")
        self.seg = seg
        rows = []
        for s, g in enumerate(self.grids):
            if not self.is_active(s):
                continue
            gx = (g.nx + self.TILE_X - 1) // self.TILE_X
            gy = (g.ny + self.seg - 1) // self.seg
            t = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="xy"), -1).reshape(-1, 2)
            rows.append(np.concatenate([np.full((len(t), 1), s), t], axis=1))
        tiles = np.concatenate(rows).astype(np.int32)
        if self.xcd:
            tiles = xcd_order(tiles)
        self.tiles = T.from_numpy(tiles.ravel()).to(self.dev)
        self.ntiles = len(tiles)
        self.graphs = {}

    def _make_slot(self, need):
        """slot[g][n]: the row of step n in point group g's record numbering, or -1."""
        T = self.torch
        slot = np.full((len(self.groups), need), -1, np.int32)
        for g, rs in enumerate(self.groups):
            rs = rs[rs < need]
            slot[g, rs] = np.arange(len(rs), dtype=np.int32)
        self.slot = T.from_numpy(slot.ravel()).to(self.dev)
        self.nslot = need
        self.graphs = {}

    def _sync_descs(self, n0, slot0=0):
        """Point the launches at descriptors for the field buffers' parity and the first snapshot slot: step n0
        must read buffer n0 & 1, and the next launch's step q writes snapshot slot slot0 + q.

        Each (parity, slot) layout is uploaded once, from pinned memory without blocking the host, into a
        device array of its own that later runs reuse (a run never waits for the GPU to drain). Captured
        graphs hold the address of the array, so with graphs one array is rewritten in stream order instead."""
        T = self.torch
        shifts = [(g._cur - (n0 & 1)) % 2 for g in self.grids]
        key = (tuple(shifts), slot0)
        if key == self.shift:
            return
        cached = None if self.use_graph else self._desc_cache.get(key)
        if cached is None:
            for d, g, sh, sn in zip(self.descs, self.grids, shifts, self.snapshots):
                for p in (0, 1):
                    b = g._fb[(p + sh) % 2]
                    # slots (ez, hx, hy) hold the grid's three components in order: (Ez, Hx, Hy) or (Hz, Ex, Ey)
                    c0, c1, c2 = g.names
                    d.ez[p], d.hx[p], d.hy[p] = b[c0].data_ptr(), b[c1].data_ptr(), b[c2].data_ptr()
                if sn is not None:
                    buf = sn[4]
                    d.snap = buf.data_ptr() + slot0 * buf[0].numel() * buf.element_size()
            raw = T.frombuffer(bytearray(b"".join(bytes(d) for d in self.descs)), dtype=T.uint8).pin_memory()
            if self.use_graph:
                cached = self.desc_dev
            else:
                cached = T.empty(raw.numel(), dtype=T.uint8, device=self.dev)
                if len(self._desc_cache) < 8:
                    self._desc_cache[key] = cached
            cached.copy_(raw, non_blocking=True)
        self.desc_dev = cached
        self.shift = key

    def _launch_k(self, nl, rd0):
        ptr = lambda t: t.data_ptr() if t is not None else None
        N.call("bz_yee2d_runk", self.desc_dev.data_ptr(), len(self.grids), self.tilesk.data_ptr(), self.ntilesk,
               self.krows, self.kwaves, self.kstep, self.step.data_ptr(), rd0, nl, ptr(self.ent_start),
               ptr(self.ents), ptr(self.stagek), self.npts, ptr(self.slot), ptr(self.grp), self.nslot, self.nrow,
               ptr(self.series), self.nfreq, ptr(self.phasor), ptr(self.dft), int(self.multi_src),
               self.grids[0].stream)

    def _launch(self, count):
        ptr = lambda t: t.data_ptr() if t is not None else None
        g0 = self.grids[0]
        N.call("bz_yee2d_hz_run" if self.pol == "Hz" else "bz_yee2d_run", self.desc_dev.data_ptr(), len(self.grids), self.tiles.data_ptr(), self.ntiles,
               self.seg, self.step.data_ptr(), count, ptr(self.pts), self.npts, ptr(self.slot), ptr(self.grp),
               self.nslot, self.nrow, ptr(self.series), self.nfreq, ptr(self.phasor), ptr(self.dft), g0.stream)

    def run(self, n0, count, coef=True):
        """Advance every grid by ``count`` steps starting at step index n0. coef=False skips recomputing the
        K-step kernel's update coefficients, for a caller that knows eps and sig are unchanged since its
        previous run (an earlier chunk of the same run)."""
        if count <= 0:
            return
        T = self.torch
        for s, ln in self.signal_lens.items():
            if self.is_active(s) and (n0 < 0 or n0 + count > ln):
                raise ValueError(f"steps {n0}..{n0 + count - 1} run past the source signal of simulation {s} "
                                 f"({ln} samples)")
        if count > getattr(self, "snap_steps", 1 << 62):
            raise ValueError(f"a run of {count} steps overflows the snapshot buffers ({self.snap_steps} slots)")
        if self.npts and n0 + count > self.nslot:
            self._make_slot(n0 + count)
        if coef:
            # the materials may have changed: the next K-step launch recomputes the coefficients, also when
            # this run is too short for one and a later run of the same materials passes coef=False
            self._coef_done = False
        self._sync_descs(n0)
        self.step.fill_(n0)
        g0 = self.grids[0]
        if g0.timing:
            e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
            e0.record()
        total = count
        if self.kstep and count >= self.kstep:
            # launches of K steps each; launch q reads the buffer holding step n0 + qK and writes the other one
            K = self.kstep
            nl = count // K
            ptr = lambda t: t.data_ptr() if t is not None else None
            # coefficients of the current materials (a caller may have written eps or sig since the last run)
            if not self._coef_done:
                N.call("bz_yee2d_coef", self.desc_dev.data_ptr(), len(self.grids),
                       max(g.ny * g.ld for g in self.grids), g0.stream)
                self._coef_done = True
            if self.use_graph:
                # one captured graph per (launch count, first buffer); the step base lives in device memory
                key = ("k", nl, n0 & 1)
                gr = self.graphs.get(key)
                if gr is None:
                    gr = T.cuda.CUDAGraph()
                    T.cuda.synchronize(self.dev)
                    with T.cuda.graph(gr):
                        self._launch_k(nl, n0 & 1)
                    self.graphs[key] = gr
                gr.replay()
            else:
                self._launch_k(nl, n0 & 1)
            for g in self._advancing():
                g._cur = (g._cur + nl) % 2
            n0, count = n0 + nl * K, count - nl * K
            if count:
                # the step counter is already there; later steps write snapshot slots from nl*K on
                self._sync_descs(n0, total - count)
        rem = count
        if self.pairs and count >= 2:
            # pairs of steps; pair q reads the buffer holding step n0 + 2q and writes the other one
            npairs = count // 2
            ptr = lambda t: t.data_ptr() if t is not None else None
            N.call("bz_yee2d_run2", self.desc_dev.data_ptr(), len(self.grids), self.tiles2.data_ptr(),
                   self.ntiles2, self.seg2, self.step.data_ptr(), n0 & 1, npairs, ptr(self.pts), ptr(self.spts),
                   self.npts, ptr(self.slot), ptr(self.grp), self.nslot, self.nrow, ptr(self.series), self.nfreq,
                   ptr(self.phasor), ptr(self.dft), g0.stream)
            for g in self._advancing():
                g._cur = (g._cur + npairs) % 2
            rem = count - 2 * npairs
            if rem:
                self._sync_descs(n0 + 2 * npairs, total - rem)     # step counter is already there
        if rem and self.use_graph and rem > 1:
            # replays of one captured GRAPH_CHUNK-step graph, then one graph for the rest: the step base lives
            # in device memory, so each replay continues where the previous one stopped
            left = rem
            while left > 1:
                c = min(left, self.GRAPH_CHUNK)
                gr = self.graphs.get(c)
                if gr is None:
                    gr = T.cuda.CUDAGraph()
                    T.cuda.synchronize(self.dev)
                    with T.cuda.graph(gr):
                        self._launch(c)
                    self.graphs[c] = gr
                gr.replay()
                left -= c
            if left:
                self._launch(left)
        elif rem:
            self._launch(rem)
        if g0.timing:
            e1.record()
            g0._events.append((e0, e1, total))
        for g in self._advancing():
            g._cur = (g._cur + rem) % 2
            if g.power is not None:
                g.power_count += total
