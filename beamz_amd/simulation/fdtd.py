"""Drop-in ``FDTD`` driver (beamz/simulation/fdtd.py:16-549) on the MI355X engine.

Same constructor, ``initialize_simulation`` / ``step`` / ``finalize_simulation``
/ ``run`` signatures, attribute names (``Ez``, ``Hx``, ``Hy`` [, ``Ex``, ``Ey``,
``Hz``], ``results``, ``monitors``, ``dt``, ``num_steps``, ``t``,
``current_step``) and result semantics as the reference. Differences, each
deliberate:

* fields are FP32 on the GPU; the field attributes return host complex128
  copies in the reference's shapes. The reference's NumPy path keeps complex
  fields (fdtd.py:151-186) and adds complex source values unreduced
  (helper.py:29-35); every update has real coefficients, so the real and the
  imaginary parts evolve independently. When a source carries a complex
  amplitude or signal, the engine runs a second real grid for the imaginary
  part (``complex_fields``, automatic by default) and the attributes, monitor
  series, spectra and power combine the two; otherwise the imaginary part is
  zero and one real grid runs. ``complex_fields=False`` keeps the real part
  only, as the reference's real-backend branch does (helper.py:29-33);
* every time step runs on the device: update, sources, monitor sampling and
  power accumulation never copy fields to the host; saved results do;
* ``run(steps=k)`` stops after k steps (the reference ignores ``steps``);
* ``live`` plotting is not performed;
* a ``GaussianSource`` passed through ``devices=`` is accepted (the reference
  raises AttributeError there because the source has no ``design``).

Extension (no reference counterpart): ``polarization="Hz"`` runs a 2D design
in the (Hz, Ex, Ey) family instead of the reference's (Ez, Hx, Hy) one. Its
update is the reference's 3D step (fdtd.py:200-310) with every d/dz = 0, kept
to Hz (ny-1, nx-1), Ex (ny, nx-1), Ey (ny-1, nx); sources add their patches or
profiles to Hz (clipped to Hz's cells), monitors sample Hz, Ex, Ey, and the
Poynting vector is S = E x H = (Ey Hz, -Ex Hz).
"""
from __future__ import annotations

import datetime
from typing import Dict, Optional, Sequence

import numpy as np

from ..const import EPS_0, LIGHT_SPEED, MU_0
from ..devices.monitors import Monitor
from ..devices.sources import GaussianSource, ModeSource
from ..helpers import check_fdtd_stability, display_status
from .backends import get_backend
from .engine import COMP3, Grid
from .meshing import BaseMeshGrid, RegularGrid, RegularGrid3D


def gaussian_weights_2d(position, width, dx, dy, nx, ny):
    """Patch and weights of a 2D GaussianSource (beamz/simulation/helper.py:96-116)."""
    cx, cy = position[0] / dx, position[1] / dy
    wx, wy = width / dx, width / dy
    hx, hy = max(1, int(round(3 * wx))), max(1, int(round(3 * wy)))
    ix, iy = int(round(cx)), int(round(cy))
    x0, x1 = max(0, ix - hx), min(nx, ix + hx + 1)
    y0, y1 = max(0, iy - hy), min(ny, iy + hy + 1)
    yy, xx = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
    w = np.exp(-((xx - cx) ** 2 / (2 * (wx ** 2 + 1e-9)) + (yy - cy) ** 2 / (2 * (wy ** 2 + 1e-9)))) / 4
    return np.stack([yy.ravel(), xx.ravel()], axis=1), w.ravel()


def gaussian_weights_3d(position, width, dx, dy, dz, nx, ny, nz, ez_nz):
    """Patch and weights of a 3D GaussianSource into Ez (beamz/simulation/helper.py:62-95)."""
    c = (position[0] / dx, position[1] / dy, position[2] / dz)
    wd = (width / dx, width / dy, width / dz)
    h = [max(1, int(round(3 * v))) for v in wd]
    ic = [int(round(v)) for v in c]
    x0, x1 = max(0, ic[0] - h[0]), min(nx, ic[0] + h[0] + 1)
    y0, y1 = max(0, ic[1] - h[1]), min(ny, ic[1] + h[1] + 1)
    z0, z1 = max(0, ic[2] - h[2]), min(nz, ic[2] + h[2] + 1)
    z1 = min(z1, ez_nz)
    zz, yy, xx = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
    w = np.exp(-((xx - c[0]) ** 2 / (2 * (wd[0] ** 2 + 1e-9)) + (yy - c[1]) ** 2 / (2 * (wd[1] ** 2 + 1e-9))
                 + (zz - c[2]) ** 2 / (2 * (wd[2] ** 2 + 1e-9))))
    zoff = max(0, min(ez_nz - 1, z0))
    zt = zz - z0 + zoff
    keep = zt < ez_nz
    idx = np.stack([zt[keep], yy[keep], xx[keep]], axis=1)
    return idx, w[keep]


def mode_source_cells(src, dx, dy, dz, nx, ny, nz, ez_nz, is_3d):
    """Affine per-cell map (mul, add) equivalent to helper.py:8-54's sequential point loop: after the loop a
    cell holds mul * (its value before) + add * signal[n]. add is complex (the profile's amplitudes); mul is 0
    where a later point's direction enforcement zeroed the cell (which also drops what earlier points added)."""
    acc = {}
    order = []

    def cell(key):
        if key not in acc:
            acc[key] = [1.0, 0j]
            order.append(key)
        return acc[key]

    enforce = getattr(src, "enforce_direction", True)
    for point in src.mode_profiles[0]:
        if len(point) == 4:
            amp, xr, yr, zr = point
        else:
            amp, xr, yr = point
            zr = 0
        x, y = int(round(xr / dx)), int(round(yr / dy))
        if is_3d:
            z = int(round(zr / dz))
            if x < 0 or x >= nx or y < 0 or y >= ny or z < 0 or z >= nz:
                continue
            zt = min(z, ez_nz - 1) if z < ez_nz else ez_nz // 2
            key = (zt, y, x)
        else:
            if x < 0 or x >= nx or y < 0 or y >= ny:
                continue
            zt = 0
            key = (y, x)
        cell(key)[1] += complex(amp)
        if enforce:
            d = src.direction
            up = None
            if d == "+x" and x > 0: up = (y, x - 1)
            elif d == "-x" and x < nx - 1: up = (y, x + 1)
            elif d == "+y" and y > 0: up = (y - 1, x)
            elif d == "-y" and y < ny - 1: up = (y + 1, x)
            elif is_3d and d == "+z" and zt > 0: up = (zt - 1, y, x)
            elif is_3d and d == "-z" and zt < ez_nz - 1: up = (zt + 1, y, x)
            if up is not None:
                if is_3d and len(up) == 2:
                    up = (zt, *up)
                cell(up)[:] = [0.0, 0j]
    idx = np.array(order, dtype=np.int64).reshape(len(order), 3 if is_3d else 2)
    mul = np.array([acc[k][0] for k in order])
    add = np.array([acc[k][1] for k in order], dtype=np.complex128)
    return idx, mul, add


class FDTD:
    def __init__(self, design, time, mesh: str = "regular", resolution: float = 0.02e-6, backend="hip",
                 backend_options=None, devices=None, *, polarization="Ez", complex_fields=None):
        provided = None
        if isinstance(design, BaseMeshGrid):
            provided = design
            if getattr(provided, "design", None) is None:
                raise ValueError("Provided mesh must expose its originating design via the `design` attribute.")
            design = provided.design
        self.design = design
        self.is_3d = bool(design.is_3d and design.depth > 0)
        if provided is not None:
            mesh_obj = provided
        elif isinstance(mesh, BaseMeshGrid):
            mesh_obj = mesh
        elif mesh == "regular":
            mesh_obj = RegularGrid3D(design, resolution) if self.is_3d else RegularGrid(design, resolution)
        else:
            raise ValueError(f"Unsupported mesh specification: {mesh!r}")
        if getattr(mesh_obj, "design", None) is not None and mesh_obj.design is not self.design:
            raise ValueError("Provided mesh was generated from a different design instance.")
        if self.is_3d and not isinstance(mesh_obj, RegularGrid3D):
            raise ValueError("3D designs require a RegularGrid3D mesh.")
        if not self.is_3d and not isinstance(mesh_obj, RegularGrid):
            raise ValueError("2D designs require a RegularGrid mesh.")
        self.mesh = mesh_obj
        self.resolution = getattr(mesh_obj, "resolution", getattr(mesh_obj, "resolution_xy", resolution))
        self.dx, self.dy = mesh_obj.dx, mesh_obj.dy
        self.dz = mesh_obj.dz if self.is_3d else None
        if self.is_3d:
            self.nz, self.ny, self.nx = mesh_obj.permittivity.shape
        else:
            self.ny, self.nx = mesh_obj.permittivity.shape
            self.nz = None
        self.epsilon_r = mesh_obj.permittivity
        self.mu_r = mesh_obj.permeability
        self.sigma = mesh_obj.conductivity
        self.backend = get_backend(name=backend, **(backend_options or {}))
        self.time = time
        self.dt = time[1] - time[0]
        self.num_steps = len(time)
        if polarization not in ("Ez", "Hz") or (self.is_3d and polarization != "Ez"):
            raise ValueError("polarization must be 'Ez' (the reference's 2D family, and 3D) or 'Hz' (2D only)")
        self.polarization = polarization
        self.grid = Grid(self.backend, self.epsilon_r, self.sigma, self.dx, self.dy, self.dt, self.dz,
                         pol=polarization)
        self.complex_fields = complex_fields      # None: a second, imaginary grid when a source is complex
        self.grid_im = None
        self._power_c = None
        self.sources = list(design.sources)
        self.monitors = list(design.monitors)
        for dev in devices or []:
            if isinstance(dev, ModeSource):
                if dev.design is None:
                    dev.design = self.design
                self.sources.append(dev)
            elif isinstance(dev, GaussianSource):
                self.sources.append(dev)
            elif isinstance(dev, Monitor):
                if dev.design is None:
                    dev.design = self.design
                self.monitors.append(dev)
            else:
                raise TypeError(f"Unsupported device type: {type(dev)}")
        names = self.grid.names
        self.results = {k: [] for k in names}
        self.results["t"] = []
        self.monitor_data = {}
        self.power_accumulated = None
        self.power_accumulation_count = 0
        self.start_time = None
        self.fig = None
        self._live = False
        self._plans = []
        self._ready = False
        self.t = 0
        self.current_step = 0

    # ---------------- field attributes ----------------
    def _host(self, name):
        a = self.grid.view(name).detach().cpu().numpy().astype(np.complex128)
        if self.grid_im is not None:
            a += 1j * self.grid_im.view(name).detach().cpu().numpy()
        return a

    def _set_field(self, name, value):
        value = np.asarray(value)
        parts = [(self.grid, np.real(value))]
        if self.grid_im is not None:
            parts.append((self.grid_im, np.imag(value)))
        for g, part in parts:
            v = part.astype(np.float32)
            tgt = g.view(name)
            if tuple(v.shape) != tuple(tgt.shape):
                raise ValueError(f"{name} must have shape {tuple(tgt.shape)}")
            tgt.copy_(self.backend.torch.from_numpy(np.ascontiguousarray(v)).to(tgt.device))

    Ez = property(lambda s: s._host("Ez"), lambda s, v: s._set_field("Ez", v))
    Hx = property(lambda s: s._host("Hx"), lambda s, v: s._set_field("Hx", v))
    Hy = property(lambda s: s._host("Hy"), lambda s, v: s._set_field("Hy", v))
    Ex = property(lambda s: s._host("Ex"), lambda s, v: s._set_field("Ex", v))
    Ey = property(lambda s: s._host("Ey"), lambda s, v: s._set_field("Ey", v))
    Hz = property(lambda s: s._host("Hz"), lambda s, v: s._set_field("Hz", v))

    def field(self, name):
        """Device (torch) view of a field component in the reference's shape; no copy. With complex fields
        this is the real part (``field_imag``: the imaginary part)."""
        return self.grid.view(name)

    def field_imag(self, name):
        """Device view of the imaginary part of a component (None when the fields are real)."""
        return None if self.grid_im is None else self.grid_im.view(name)

    @property
    def is_complex_backend(self):
        """True when the fields carry an imaginary part (the reference's complex128 fields)."""
        return self.grid_im is not None

    @is_complex_backend.setter
    def is_complex_backend(self, value):
        pass                            # set by the reference's constructor; here it follows the sources

    def _wants_complex(self):
        if self.complex_fields is not None:
            return bool(self.complex_fields)
        for src in self.sources:
            sig = np.asarray(src.signal)
            if np.iscomplexobj(sig) and np.any(np.imag(sig[: self.num_steps]) != 0):
                return True
            if isinstance(src, ModeSource):
                amps = np.asarray([p[0] for p in src.mode_profiles[0]])
                if np.iscomplexobj(amps) and np.any(np.imag(amps) != 0):
                    return True
        return False

    def _grids(self):
        return [self.grid] if self.grid_im is None else [self.grid, self.grid_im]

    # ---------------- setup ----------------
    def _time_table(self):
        t = np.empty(self.num_steps + 1)
        acc = 0
        for n in range(self.num_steps + 1):
            t[n] = acc
            acc += self.dt
        return t

    def _plan_sources(self):
        """Point sources of every design source on the real grid and, with complex fields, on the imaginary
        one. A source adds add * signal[n] (complex add and signal) after mul * field: the real grid gets
        Re(add) Re(s) - Im(add) Im(s), the imaginary grid Re(add) Im(s) + Im(add) Re(s), each product term a
        point source of its own with the direction-enforcing mul on the first."""
        for g in self._grids():
            g.sources = []
            g.invalidate()
        ez_nz = self.nz - 1 if self.is_3d else None
        target = self.grid.names[0] if not self.is_3d else "Ez"
        # the (Hz, Ex, Ey) family injects into Hz, whose cells are (ny-1, nx-1)
        bx, by = (self.nx - 1, self.ny - 1) if target == "Hz" else (self.nx, self.ny)
        for src in self.sources:
            sig = np.asarray(src.signal)
            if sig.ndim == 0 or len(sig) < self.num_steps:
                raise ValueError("source signal must hold at least one value per time step")
            if isinstance(src, ModeSource):
                idx, mul, add = mode_source_cells(src, self.dx, self.dy, self.dz, bx, by, self.nz, ez_nz,
                                                  self.is_3d)
            else:
                if self.is_3d:
                    idx, w = gaussian_weights_3d(src.position, src.width, self.dx, self.dy, self.dz, self.nx, self.ny,
                                                 self.nz, ez_nz)
                else:
                    idx, w = gaussian_weights_2d(src.position, src.width, self.dx, self.dy, self.nx, self.ny)
                    keep = (idx[:, 0] < by) & (idx[:, 1] < bx)
                    idx, w = idx[keep], w[keep]
                mul, add = None, np.asarray(w, dtype=np.complex128)
            if not len(idx):
                continue
            s_re, s_im = np.real(sig).astype(np.float64), np.imag(sig).astype(np.float64)
            a_re, a_im = np.real(add), np.imag(add)
            terms = {id(self.grid): [(a_re, s_re), (-a_im, s_im)]}
            if self.grid_im is not None:
                terms[id(self.grid_im)] = [(a_im, s_re), (a_re, s_im)]
            for g in self._grids():
                first = True
                for a, sg in terms[id(g)]:
                    # a term that adds nothing is kept when it carries the zeroing mul
                    if not first and (not np.any(a) or not np.any(sg)):
                        continue
                    if first and mul is None and (not np.any(a) or not np.any(sg)):
                        continue
                    g.add_point_source(target, idx, a, mul if first else None, sg)
                    first = False

    def _monitor_samples(self, mon):
        """(comp indices, offsets, layout) of one monitor, following monitors.py:204-273, 324-379."""
        g = self.grid
        shp = g.shapes
        if not self.is_3d:
            pts = mon.get_grid_points_2d(self.dx, self.dy)
            # the second component (Hx, or Ex) is never written at column nx-1: a zero to read
            sentinel = (1, g.offset(0, 0, self.nx - 1))
            comps, offs = [], []
            for ci, name in enumerate(g.names):
                sy, sx = shp[name]
                for x, y in pts:
                    if 0 <= y < sy and 0 <= x < sx:
                        comps.append(ci)
                        offs.append(g.offset(0, int(y), int(x)))
                    else:
                        comps.append(sentinel[0])
                        offs.append(sentinel[1])
            return comps, offs, ("line", len(pts))
        sl = mon.get_grid_slice_3d(self.dx, self.dy, self.dz)
        normal = mon.plane_normal
        regions = {}
        for name in COMP3:
            a0, a1, a2 = shp[name]
            if normal == "z":
                zi = min(max(0, sl[2]), a0 - 1)
                ys = sl[0].start or 0; ye = sl[0].stop or a1
                xs = sl[1].start or 0; xe = sl[1].stop or a2
                ys = max(0, min(ys, a1 - 1)); ye = max(ys, min(ye, a1))
                xs = max(0, min(xs, a2 - 1)); xe = max(xs, min(xe, a2))
                regions[name] = ("z", zi, ys, ye - ys, xs, xe - xs)
            elif normal == "y":
                yi = min(max(0, sl[0]), a1 - 1)
                xs, xe = sl[1].start or 0, sl[1].stop or a2
                zs, ze = sl[2].start or 0, sl[2].stop or a0
                xs = max(0, min(xs, a2 - 1)); xe = max(xs, min(xe, a2))
                zs = max(0, min(zs, a0 - 1)); ze = max(zs, min(ze, a0))
                regions[name] = ("y", yi, zs, ze - zs, xs, xe - xs)
            else:
                xi = min(max(0, sl[1]), a2 - 1)
                ys, ye = sl[0].start or 0, sl[0].stop or a1
                zs, ze = sl[2].start or 0, sl[2].stop or a0
                ys = max(0, min(ys, a1 - 1)); ye = max(ys, min(ye, a1))
                zs = max(0, min(zs, a0 - 1)); ze = max(zs, min(ze, a0))
                regions[name] = ("x", xi, zs, ze - zs, ys, ye - ys)
        m0 = min(r[3] for r in regions.values())
        m1 = min(r[5] for r in regions.values())
        comps, offs = [], []
        for ci, name in enumerate(COMP3):
            kind, fixed, s0, _, s1, _ = regions[name]
            a, b = np.meshgrid(np.arange(s0, s0 + m0), np.arange(s1, s1 + m1), indexing="ij")
            if kind == "z":
                o = (fixed * self.ny + a) * g.ld + b
            elif kind == "y":
                o = (a * self.ny + fixed) * g.ld + b
            else:
                o = (a * self.ny + b) * g.ld + fixed
            comps += [ci] * o.size
            offs += o.ravel().tolist()
        return comps, offs, ("plane", (m0, m1))

    def _plan_monitors(self, t_all):
        g = self.grid
        for gg in self._grids():
            gg.monitors = []
            gg.invalidate()
        self._plans = []
        for mon in self.monitors:
            if not self.is_3d and set(mon._fields) != set(g.names) | {"t"}:
                mon._fields = {k: [] for k in g.names}
                mon._fields["t"] = []
            comps, offs, layout = self._monitor_samples(mon)
            last = mon.last_record_step
            steps = []
            for n in range(self.num_steps):
                if n - last >= mon.record_interval:
                    steps.append(n)
                    last = n
            omegas = None if mon.dft_frequencies is None else 2 * np.pi * mon.dft_frequencies
            keep = bool(mon.should_record_fields or mon.accumulate_power)
            plan = g.add_gather(comps, offs, steps, t_all, omegas, keep_series=keep)
            # the imaginary grid samples the same points; its series and spectra are the imaginary parts
            plan.im = None if self.grid_im is None else self.grid_im.add_gather(comps, offs, steps, t_all, omegas,
                                                                                keep_series=keep)
            plan.monitor, plan.layout, plan.t_all = mon, layout, t_all
            plan.fetched = 0
            mon._pull = self._make_pull(plan)
            self._plans.append(plan)

    def _make_pull(self, plan):
        def pull(mon):
            self._materialize(plan)
        return pull

    def _materialize(self, plan):
        """Copy a monitor's device samples into the reference's host structures."""
        mon = plan.monitor
        done = [s for s in plan.record_steps if s < self.current_step]
        nrec = len(done)
        new = slice(plan.fetched, nrec)
        names = self.grid.names
        if nrec > plan.fetched and plan.series is not None:
            ser = plan.series[new].detach().cpu().numpy().astype(np.float64)
            if plan.im is not None:
                ser = ser + 1j * plan.im.series[new].detach().cpu().numpy()
            ts = [float(plan.t_all[s]) if s > 0 else 0 for s in done[new]]
            if self.is_3d:
                m0, m1 = plan.layout[1]
                blk = ser.reshape(ser.shape[0], 6, m0, m1)
                comp = {k: blk[:, q] for q, k in enumerate(names)}
            else:
                npts = plan.layout[1]
                blk = ser.reshape(ser.shape[0], 3, npts)
                comp = {k: blk[:, q] for q, k in enumerate(names)}
            if mon.should_record_fields:
                for k in names:
                    mon._fields[k].extend(list(comp[k]) if self.is_3d else [row.tolist() for row in comp[k]])
                mon._fields["t"].extend(ts)
            if mon.accumulate_power:
                if self.is_3d:
                    ex, ey, ez, hx, hy, hz = (comp[k] for k in names)
                    sx, sy, sz = ey * hz - ez * hy, ez * hx - ex * hz, ex * hy - ey * hx
                    pm = np.sqrt(sx ** 2 + sy ** 2 + sz ** 2)
                    tot = pm.reshape(pm.shape[0], -1).sum(axis=1)
                elif self.polarization == "Hz":
                    hz, ex, ey = comp["Hz"], comp["Ex"], comp["Ey"]
                    pm = np.sqrt((ey * hz) ** 2 + (-ex * hz) ** 2)
                    tot = pm.sum(axis=1)
                else:
                    ez, hx, hy = comp["Ez"], comp["Hx"], comp["Hy"]
                    pm = np.sqrt((-ez * hy) ** 2 + (ez * hx) ** 2)
                    tot = pm.sum(axis=1)
                acc = pm.sum(axis=0)
                mon._power_accumulated = acc if mon._power_accumulated is None else mon._power_accumulated + acc
                mon._power_history.extend(tot.tolist())
                mon._power_timestamps.extend(float(t) for t in ts)
                mon.power_accumulation_count += nrec - plan.fetched
            m = mon.max_history_steps
            if m is not None:
                for k in mon._fields:
                    if len(mon._fields[k]) > m:
                        mon._fields[k] = mon._fields[k][len(mon._fields[k]) - m:]
                if len(mon._power_history) > m:
                    mon._power_history = mon._power_history[len(mon._power_history) - m:]
                    mon._power_timestamps = mon._power_timestamps[len(mon._power_timestamps) - m:]
        if nrec:
            mon.last_record_step = int(done[-1])
        plan.fetched = nrec
        if plan.nfreq:
            d = plan.dft.detach().cpu().numpy()
            z = d[..., 0] + 1j * d[..., 1]
            if plan.im is not None:
                d = plan.im.dft.detach().cpu().numpy()
                z = z + 1j * (d[..., 0] + 1j * d[..., 1])      # DFT(re) + i DFT(im)
            if self.is_3d:
                m0, m1 = plan.layout[1]
                z = z.reshape(plan.nfreq, 6, m0, m1)
            else:
                z = z.reshape(plan.nfreq, 3, plan.layout[1])
            mon.dft = {k: z[:, q] for q, k in enumerate(names)}
        if self.current_step < self.num_steps:
            mon._pull = self._make_pull(plan)

    def initialize_simulation(self, save=True, live=True, axis_scale=[-1, 1], save_animation=False,
                              animation_filename="fdtd_animation.mp4", clean_visualization=True, save_fields=None,
                              decimate_save=1, accumulate_power=False, save_memory_mode=False,
                              fields_to_cache: Optional[Sequence[str]] = None):
        if save_fields is None:
            save_fields = list(self.grid.names)
        self._cache_fields = list(fields_to_cache) if fields_to_cache else []
        self._cache_frequency = 1 if self._cache_fields else None
        self.start_time = datetime.datetime.now()
        self.t = 0
        self.current_step = 0
        self._total_steps = self.num_steps
        self._save_results = save
        self._save_fields = save_fields
        self._decimate_save = decimate_save
        self._live = live
        self._axis_scale = axis_scale
        for k in list(self.results.keys()):
            self.results[k] = []
        self.save_memory_mode = save_memory_mode
        self.accumulate_power = accumulate_power
        n_max = np.sqrt(np.max(self.epsilon_r))
        ok, courant, lim = check_fdtd_stability(dt=self.dt, dx=self.dx, dy=self.dy, n_max=n_max, safety_factor=1.0)
        if not ok:
            display_status(f"Simulation may be unstable! Courant number = {courant:.3f} > {lim:.3f}", "warning")
        self._effective_save_freq = 1 * decimate_save
        if self._cache_frequency is None:
            self._cache_frequency = self._effective_save_freq
        if self._wants_complex():
            if self.grid_im is None:
                self.grid_im = Grid(self.backend, self.epsilon_r, self.sigma, self.dx, self.dy, self.dt, self.dz,
                                    pol=self.polarization)
                self._copy_settings(self.grid, self.grid_im)
            # the imaginary grid shares the real grid's materials, also when they were edited on the device
            self.grid_im.eps.copy_(self.grid.eps)
            self.grid_im.sig.copy_(self.grid.sig)
        else:
            self.grid_im = None
        for g in self._grids():
            g.reset_fields()
            g.power = None
            g.invalidate()
        self.power_accumulated = None
        self.power_accumulation_count = 0
        self._power_c = None
        if accumulate_power:
            if self.grid_im is None:
                self.grid.enable_power()
            else:
                # Re(E x conj H) needs both parts of each step: accumulated here after both grids stepped
                self._power_c = self.backend.torch.zeros(self.grid.box, dtype=self.backend.torch.float64,
                                                         device=self.backend.dev)
        self._t_all = self._time_table()
        self._plan_sources()
        self._plan_monitors(self._t_all)
        self._ready = True

    def _needs_host(self, n):
        if not self._save_results:
            return False
        full = (not self.save_memory_mode) and (n % self._effective_save_freq == 0 or n == self.num_steps - 1)
        cache = bool(self._cache_fields) and (n % self._cache_frequency == 0 or n == self.num_steps - 1)
        return full or cache

    def _save(self, n):
        """helper.py:189-229."""
        full = (not self.save_memory_mode) and (n % self._effective_save_freq == 0 or n == self.num_steps - 1)
        cache = bool(self._cache_fields) and (n % self._cache_frequency == 0 or n == self.num_steps - 1)
        if not (self._save_results and (full or cache)):
            return
        self.results.setdefault("t", []).append(self.t)
        names = list(self._save_fields) if full else []
        if cache:
            names += [f for f in self._cache_fields if f not in names]
        for name in names:
            a = self._host(name)
            if name not in self._cache_fields:
                a = np.abs(a)
            self.results.setdefault(name, []).append(a)

    def step(self) -> bool:
        if not self._ready:
            self.initialize_simulation(live=False)
        if self.current_step >= self.num_steps:
            return False
        n = self.current_step
        self._run_grids(n, 1)
        self._save(n)
        self.t += self.dt
        self.current_step += 1
        return True

    def _advance(self, count):
        """Run ``count`` steps, batching device work between host saves."""
        end = min(self.num_steps, self.current_step + count)
        while self.current_step < end:
            n = self.current_step
            stop = n + 1
            if not self._needs_host(n):
                while stop < end and not self._needs_host(stop):
                    stop += 1
                self._run_grids(n, stop - n)
                for _ in range(stop - n):
                    self.t += self.dt
                self.current_step = stop
            else:
                self.step()

    @staticmethod
    def _copy_settings(src, dst):
        for k in ("kernel", "seg2d", "graph2d", "kstep2d", "krows2d", "kwaves2d", "kseg", "rows", "tb2", "kseg2",
                  "rows2", "waves2"):
            setattr(dst, k, getattr(src, k))

    def _run_grids(self, n, count):
        """Steps n .. n+count-1 on the real grid and, with complex fields, on the imaginary one (with grid
        power: step by step, accumulating |S| of the combined fields after each)."""
        if self.grid_im is None:
            self.grid.run(n, count)
            return
        self._copy_settings(self.grid, self.grid_im)
        if self._power_c is None:
            self.grid.run(n, count)
            self.grid_im.run(n, count)
            return
        from .. import _native as N
        g, h = self.grid, self.grid_im
        for q in range(n, n + count):
            g.run(q, 1)
            h.run(q, 1)
            if self.is_3d:
                N.call("bz_grid_power_3d_cplx", self._power_c.data_ptr(), N.ptrs6([g.f[k].data_ptr() for k in COMP3]),
                       N.ptrs6([h.f[k].data_ptr() for k in COMP3]), self.nx, self.ny, self.nz, g.ld, g.stream)
            else:
                if self.polarization != "Ez":
                    raise NotImplementedError("grid power of complex (Hz, Ex, Ey) fields")
                N.call("bz_grid_power_2d_cplx", self._power_c.data_ptr(), g.f["Ez"].data_ptr(), g.f["Hx"].data_ptr(),
                       g.f["Hy"].data_ptr(), h.f["Ez"].data_ptr(), h.f["Hx"].data_ptr(), h.f["Hy"].data_ptr(),
                       self.nx, self.ny, g.ld, g.stream)
            self.power_accumulation_count += 1

    def finalize_simulation(self):
        if self._power_c is not None:
            p = self._power_c.detach().cpu().numpy()
            if self.is_3d:
                p = p[: self.nz - 1, : self.ny - 1, : self.nx - 1]
            else:
                p = p[:, : self.nx]
            c = self.power_accumulation_count
            self.power_accumulated = p / c if c else p
        elif self.grid.power is not None:
            self.power_accumulation_count = self.grid.power_count
            p = self.grid.power.detach().cpu().numpy()
            if self.is_3d:
                p = p[: self.nz - 1, : self.ny - 1, : self.nx - 1]
            else:
                p = p[:, : self.nx]
            self.power_accumulated = p / self.power_accumulation_count if self.power_accumulation_count else p
        for plan in self._plans:
            self._materialize(plan)
        objectives: Dict[str, float] = {}
        for q, mon in enumerate(self.monitors):
            v = mon.evaluate_objective() if hasattr(mon, "evaluate_objective") else None
            if v is None:
                continue
            objectives[getattr(mon, "name", None) or f"monitor_{q}"] = v
        if objectives:
            self.results["objectives"] = objectives
        else:
            self.results.pop("objectives", None)
        self.last_objectives = objectives
        return self.results

    def run(self, steps: Optional[int] = None, save=True, live=True, axis_scale=[-1, 1], save_animation=False,
            animation_filename="fdtd_animation.mp4", clean_visualization=True, save_fields=None, decimate_save=1,
            accumulate_power=False, save_memory_mode=False, fields_to_cache: Optional[Sequence[str]] = None) -> Dict:
        self.initialize_simulation(save=save, live=live, axis_scale=axis_scale, save_animation=save_animation,
                                   animation_filename=animation_filename, clean_visualization=clean_visualization,
                                   save_fields=save_fields, decimate_save=decimate_save,
                                   accumulate_power=accumulate_power, save_memory_mode=save_memory_mode,
                                   fields_to_cache=fields_to_cache)
        self._advance(self.num_steps if steps is None else int(steps))
        bad = self.check_fields()
        if bad:
            display_status(f"non-finite field values after {self.current_step} steps: {bad}", "warning")
        return self.finalize_simulation()

    def check_fields(self):
        """Numerical watchdog (no reference counterpart): {component: number of NaN or infinite entries} of
        the current fields, empty when every entry is finite. One pass over the device fields."""
        from .. import _native as N
        T = self.backend.torch
        names = list(self.grid.names)
        count = T.zeros(len(names), dtype=T.int32, device=self.backend.dev)
        for g in self._grids():
            for q, k in enumerate(names):
                t = g.f[k]
                N.call("bz_nonfinite_count", t.data_ptr(), t.numel(), count.data_ptr() + 4 * q, g.stream)
        c = count.cpu().numpy()
        return {k: int(v) for k, v in zip(names, c) if v}

    def synchronize(self):
        self.backend.torch.cuda.synchronize(self.backend.dev)

    def estimate_memory_usage(self, time_steps=None, save_fields=None):
        """helper.py:258-302 (8 bytes per saved value, as the reference assumes)."""
        steps = self.num_steps if time_steps is None else time_steps
        names = save_fields or list(self.grid.names)
        sizes = {k: int(np.prod(v)) * 8 for k, v in self.grid.shapes.items()}
        total = steps * 8 + sum(sizes.get(k, 0) * steps for k in names)
        mb = 1024 * 1024
        return {"Single timestep": {**{k: sizes.get(k, 0) / mb for k in names},
                                    "Total": sum(sizes.get(k, 0) for k in names) / mb},
                "Full simulation": {"Total memory (MB)": total / mb, "Total memory (GB)": total / mb / 1024,
                                    "Time steps": steps}}
