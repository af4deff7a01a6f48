"""Field monitors (beamz/devices/monitors.py:7-453).

The reference samples fields on the host every step. Here the engine gathers
the monitor's samples on the GPU every step into a device time-series buffer
and, if frequencies are requested, into FP64 DFT accumulators; ``fields``,
``power_history``, ``power_accumulated`` and ``power_timestamps`` are filled
from the device buffers when read (same layout as the reference: one list per
recorded step). Field values are the real fields (the reference's complex128
fields have zero imaginary part for real source signals).

New (no reference counterpart): ``dft_frequencies`` (Hz). After a run,
``monitor.dft[name]`` holds sum_n f_n exp(-i w t_n) for each recorded
component, shape (n_freq, n_points) in 2D or (n_freq, *plane_shape) in 3D.

Plane monitors: z-normal planes follow the reference's slicing
(monitors.py:222-273, 324-337). x- and y-normal planes are sampled as true
planes here (the reference copies the whole volume for them); they are the
flux planes for 3D ports.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np


class Monitor:
    def __init__(self, design=None, start=(0, 0), end=None, plane_normal=None, plane_position=0, size=None,
                 record_fields=True, accumulate_power=True, live_update=False, record_interval=1,
                 max_history_steps=None, objective_function: Optional[Callable[["Monitor"], float]] = None,
                 name: Optional[str] = None, dft_frequencies=None):
        self.design = design
        self.should_record_fields = record_fields
        self.accumulate_power = accumulate_power
        self.live_update = live_update
        self.record_interval = record_interval
        self.max_history_steps = max_history_steps
        self.objective_function = objective_function
        self.objective_value = None
        self.name = name
        self.dft_frequencies = None if dft_frequencies is None else np.atleast_1d(np.asarray(dft_frequencies, float))
        self.dft = {}
        self.is_3d = self._determine_3d_mode(start, end, design)
        names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz") if self.is_3d else ("Ez", "Hx", "Hy")
        self._fields = {k: [] for k in names}
        self._fields["t"] = []
        self._power_accumulated = None
        self._power_history = []
        self._power_timestamps = []
        self.energy_history = []
        self.power_accumulation_count = 0
        self.step_count = 0
        self.last_record_step = -1
        self._pull = None          # set by the engine: callable that materialises device buffers
        if self.is_3d:
            self._init_3d_monitor(start, end, plane_normal, plane_position, size)
        else:
            self._init_2d_monitor(start, end)

    # ---- host views filled from the device (lazy) ----
    def _materialize(self):
        if self._pull is not None:
            pull, self._pull = self._pull, None
            pull(self)

    @property
    def fields(self):
        self._materialize()
        return self._fields

    @property
    def power_history(self):
        self._materialize()
        return self._power_history

    @property
    def power_timestamps(self):
        self._materialize()
        return self._power_timestamps

    @property
    def power_accumulated(self):
        self._materialize()
        return self._power_accumulated

    # assignment replaces the host copy (load_data does this, as in the reference)
    @fields.setter
    def fields(self, value):
        self._pull = None
        self._fields = value

    @power_history.setter
    def power_history(self, value):
        self._pull = None
        self._power_history = value

    @power_timestamps.setter
    def power_timestamps(self, value):
        self._pull = None
        self._power_timestamps = value

    # ---- geometry (monitors.py:76-220) ----
    def _determine_3d_mode(self, start, end, design):
        if end is not None and len(end) == 3:
            return True
        if len(start) == 3:
            return True
        if end is not None and len(start) == 2 and len(end) == 2:
            return False
        return bool(design and getattr(design, "is_3d", False))

    def _init_2d_monitor(self, start, end):
        if end is None:
            end = start
        self.start, self.end = start, end
        self.position = ((start[0] + end[0]) / 2, (start[1] + end[1]) / 2)
        self.monitor_type = "line"

    def _init_3d_monitor(self, start, end, plane_normal, plane_position, size):
        if len(start) == 2:
            start = (start[0], start[1], 0.0)
        self.start = start
        if end is not None:
            if len(end) == 2:
                end = (end[0], end[1], start[2])
            self.end = end
            self.size = (abs(end[0] - start[0]), abs(end[1] - start[1]))
            self.plane_position = start[2]
            self.plane_normal = "z"
            self.start = (min(start[0], end[0]), min(start[1], end[1]), start[2])
        else:
            self.end = None
            self.plane_normal = plane_normal or "z"
            self.plane_position = plane_position
            if size is None:
                d = self.design
                if d:
                    size = {"z": (d.width, d.height), "y": (d.width, d.depth or d.width)}.get(
                        self.plane_normal, (d.height, d.depth or d.height))
                else:
                    size = (1e-6, 1e-6)
            self.size = size
        self.monitor_type = "plane"
        s, z = self.start, self.plane_position
        if self.plane_normal == "z" or self.end is not None:
            self.position = (s[0] + self.size[0] / 2, s[1] + self.size[1] / 2, z)
        elif self.plane_normal == "y":
            self.position = (s[0] + self.size[0] / 2, z, s[2] + self.size[1] / 2)
        else:
            self.position = (z, s[1] + self.size[0] / 2, s[2] + self.size[1] / 2)

    def get_grid_points_2d(self, dx, dy):
        """monitors.py:204-220: integer points from numpy.linspace(..., dtype=int)."""
        sx, sy = int(round(self.start[0] / dx)), int(round(self.start[1] / dy))
        ex, ey = int(round(self.end[0] / dx)), int(round(self.end[1] / dy))
        n = (abs(ex - sx) if abs(ex - sx) > abs(ey - sy) else abs(ey - sy)) + 1
        return list(zip(np.linspace(sx, ex, n, dtype=int), np.linspace(sy, ey, n, dtype=int)))

    def get_grid_slice_3d(self, dx, dy, dz, field_shape=None):
        """monitors.py:222-273."""
        d = self.design
        bnx = max(1, int(round(getattr(d, "width", 0.0) / dx)))
        bny = max(1, int(round(getattr(d, "height", 0.0) / dy)))
        bnz = max(1, int(round((getattr(d, "depth", 0.0) or 0.0) / dz)))
        clamp = lambda v, hi: max(0, min(v, hi))
        st, sz = self.start, self.size
        if self.plane_normal == "z":
            zi = clamp(int(round(self.plane_position / dz)), bnz - 1)
            return (slice(clamp(int(round(st[1] / dy)), bny - 1), clamp(int(round((st[1] + sz[1]) / dy)), bny)),
                    slice(clamp(int(round(st[0] / dx)), bnx - 1), clamp(int(round((st[0] + sz[0]) / dx)), bnx)), zi)
        if self.plane_normal == "y":
            yi = clamp(int(round(self.plane_position / dy)), bny - 1)
            return (yi, slice(clamp(int(round(st[0] / dx)), bnx - 1), clamp(int(round((st[0] + sz[0]) / dx)), bnx)),
                    slice(clamp(int(round(st[2] / dz)), bnz - 1), clamp(int(round((st[2] + sz[1]) / dz)), bnz)))
        xi = clamp(int(round(self.plane_position / dx)), bnx - 1)
        return (slice(clamp(int(round(st[1] / dy)), bny - 1), clamp(int(round((st[1] + sz[0]) / dy)), bny)), xi,
                slice(clamp(int(round(st[2] / dz)), bnz - 1), clamp(int(round((st[2] + sz[1]) / dz)), bnz)))

    def should_record(self, step):
        return (step - self.last_record_step) >= self.record_interval

    def dft_flux(self):
        """Net spectral power through the monitor per DFT frequency (new, no reference counterpart).

        Sums Re(S . n) over the monitor points from the DFT fields, with the
        reference's own Poynting conventions: in 2D (monitors.py:407-421)
        Sx = -Ez Hy and Sy = Ez Hx, here with conj(H) on the DFT amplitudes
        (the (Hz, Ex, Ey) family: Sx = Ey Hz, Sy = -Ex Hz); a vertical line is
        x-normal, a horizontal one y-normal. In 3D, the plane's normal
        component of E x conj(H). Units are field-product sums (no cell area
        factor): ratios between monitors of one grid are the port
        transmissions.

        The products pair the E and H samples of the same array index, as the
        reference's own power (monitors.py:426-434) does, without moving them
        to a common point of the Yee cell: the components sit up to half a cell
        apart in space, and H half a step apart in time. For a port plane
        across a guided mode this changes the flux by a factor close to 1 that
        is the same at every port of one grid, so port ratios keep it only to
        that order.
        """
        self._materialize()
        if not self.dft:
            raise ValueError("dft_flux needs dft_frequencies and a finished run")
        D = self.dft
        if not self.is_3d:
            vertical = abs(self.end[0] - self.start[0]) <= abs(self.end[1] - self.start[1])
            if "Hz" in D:       # the (Hz, Ex, Ey) family: S = E x conj(H) = (Ey conj(Hz), -Ex conj(Hz))
                s = D["Ey"] * np.conj(D["Hz"]) if vertical else -D["Ex"] * np.conj(D["Hz"])
            else:
                s = -D["Ez"] * np.conj(D["Hy"]) if vertical else D["Ez"] * np.conj(D["Hx"])
            return np.real(s).reshape(s.shape[0], -1).sum(axis=1)
        n = self.plane_normal if self.end is None else "z"
        a, b = {"x": ("y", "z"), "y": ("z", "x"), "z": ("x", "y")}[n]
        s = D["E" + a] * np.conj(D["H" + b]) - D["E" + b] * np.conj(D["H" + a])
        return np.real(s).reshape(s.shape[0], -1).sum(axis=1)

    def evaluate_objective(self) -> Optional[float]:
        """monitors.py:58-74."""
        if self.objective_function is None:
            return None
        try:
            v = self.objective_function(self)
        except Exception as exc:  # the reference reports and continues
            print(f"Warning: monitor objective evaluation failed: {exc}")
            return None
        if v is None:
            return None
        try:
            self.objective_value = float(v)
        except (TypeError, ValueError):
            print(f"Warning: monitor objective returned non-numeric value: {v}")
            return None
        return self.objective_value

    # ---- recorded-data access and files (monitors.py:549-589, 787-837) ----
    def get_field_statistics(self):
        """monitors.py:549-567: record count, time span, mean and peak power, and the monitor's geometry."""
        f = self.fields
        if not f["t"]:
            return {}
        ph = self.power_history
        stats = {
            "total_records": len(f["t"]),
            "time_span": f["t"][-1] - f["t"][0] if len(f["t"]) > 1 else 0,
            "avg_power": np.mean(ph) if ph else 0,
            "max_power": np.max(ph) if ph else 0,
            "monitor_type": self.monitor_type,
            "is_3d": self.is_3d,
        }
        if self.is_3d:
            stats["plane_normal"] = self.plane_normal
            stats["plane_position"] = self.plane_position
            stats["plane_size"] = self.size
        else:
            stats["line_start"] = self.start
            stats["line_end"] = self.end
        return stats

    def get_field_at_time(self, field="Ez", time_value=None, time_index=None):
        """monitors.py:787-814: the record at time_index, else the record nearest time_value, else the last."""
        f = self.fields
        if not f["t"] or field not in f:
            return None
        if time_index is not None:
            return f[field][time_index] if 0 <= time_index < len(f[field]) else None
        if time_value is not None:
            return f[field][int(np.argmin(np.abs(np.array(f["t"]) - time_value)))]
        return f[field][-1] if f[field] else None

    def get_power_statistics(self):
        """monitors.py:816-833."""
        ph = self.power_history
        if not ph:
            return {}
        a = np.array(ph)
        m = np.mean(a)
        return {"mean_power": m, "max_power": np.max(a), "min_power": np.min(a), "std_power": np.std(a),
                "total_energy": np.sum(a), "peak_to_average_ratio": np.max(a) / m if m > 0 else 0}

    def save_data(self, filename, format="npz"):
        """Write the recorded data to an .npz file (monitors.py:569-579 names the same contents).

        The reference stores ``fields`` and ``monitor_info`` as pickled dicts. Here every entry is a plain
        array, so the file loads without unpickling: ``fields/<name>`` holds the stacked records of each
        recorded component (records first), ``power_history`` and ``power_timestamps`` the power series,
        and ``monitor_info`` a JSON string with the monitor type and dimensionality."""
        import json
        if format != "npz":
            raise ValueError(f"Unsupported format: {format}")
        f = self.fields
        arrays = {f"fields/{k}": np.asarray(v) for k, v in f.items()}
        np.savez(filename, power_history=np.asarray(self.power_history, dtype=float),
                 power_timestamps=np.asarray(self.power_timestamps),
                 monitor_info=np.array(json.dumps({"type": self.monitor_type, "is_3d": bool(self.is_3d)})),
                 **arrays)

    def load_data(self, filename, allow_pickle=False):
        """Read a file written by save_data (monitors.py:581-589) back into fields / power_history /
        power_timestamps; timestamps default to 0..len-1 when absent, as in the reference. A file in the
        reference's own layout stores ``fields`` pickled: it is read only with allow_pickle=True, which
        runs code from the file, so pass it only for files you trust."""
        with np.load(filename, allow_pickle=allow_pickle) as data:
            keys = set(data.files)
            if "fields" in keys:
                try:
                    fields = data["fields"].item()
                except ValueError as exc:
                    raise ValueError(f"{filename} stores pickled fields (the reference's layout); "
                                     "load it with allow_pickle=True if you trust the file") from exc
            else:
                fields = {k[len("fields/"):]: list(data[k]) for k in data.files if k.startswith("fields/")}
            self.fields = fields
            self.power_history = list(data["power_history"])
            self.power_timestamps = (list(data["power_timestamps"]) if "power_timestamps" in keys
                                     else list(range(len(self.power_history))))

    def __str__(self):
        stats = self.get_field_statistics()
        return f"Monitor: {stats['monitor_type']} ({'3D' if stats['is_3d'] else '2D'}), {stats['total_records']} records"

    def copy(self):
        m = Monitor.__new__(Monitor)
        m.__dict__.update({k: v for k, v in self.__dict__.items()})
        m._fields = {k: [] for k in self._fields}
        m._power_history, m._power_timestamps, m._power_accumulated, m._pull, m.dft = [], [], None, None, {}
        return m
