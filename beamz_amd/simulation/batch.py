"""Many independent 2D simulations advanced together on one GPU.

New API (the reference runs one ``FDTD`` at a time, fdtd.py:411-429 per step):
``FDTDBatch`` takes ordinary 2D ``FDTD`` objects (each built from its own
design, time axis, source and monitors exactly as for a single run) and
advances all of them through one work list of the fused 2D engine. By default
(``kstep=None``: on unless a member records field power; the region shape
follows the batch's cell count) one launch of ``bz_yee2d_runk`` advances every
member K steps: each simulation is cut into regions of 64 columns x
(rows x waves) rows whose inner cells are stored after K steps held in
registers. ``kstep=0`` falls back to one or two steps per launch
(``bz_yee2d_run`` / the pair kernel) on tiles of 62 columns x ``seg`` rows. The
monitor samples and FP64 DFT sums of every member are gathered into one table.
Each member's step is bit-for-bit the step it takes when run alone.

Members may differ in everything a single run may: design, grid size,
resolution, time step and time axis length, source and signal (wavelength),
and monitors with their own record steps and DFT frequencies (each distinct
schedule and frequency list is a point group with its own slot row and phasor
block). A member whose time axis is shorter stops at its end: the engine then
rebuilds its tile lists without it and the others go on. A member may have
several sources on disjoint cells (one source box per member, each cell
reading its own source's signal). ``autotune``
picks the kernel shape (K, rows, waves; or the tile height without K steps) by
timing a few steps of each candidate on the batch itself (HIP events), then
restores the initial state.
"""
from __future__ import annotations

import numpy as np

from .engine import Fused2D


class FDTDBatch:
    SEGS = Fused2D.SEGS

    def __init__(self, sims, seg=None, graph=False, chunk=96, pairs=None, seg2=None, kstep=None, krows=None,
                 kwaves=None):
        if not sims:
            raise ValueError("FDTDBatch needs at least one simulation")
        for s in sims:
            if s.is_3d:
                raise ValueError("FDTDBatch runs 2D simulations only")
        self.sims = list(sims)
        self.lengths = np.array([len(s.time) for s in sims], np.int64)   # steps of each member's time axis
        self.seg, self.graph, self.chunk = seg, bool(graph), int(chunk)
        self.pairs, self.seg2 = pairs, seg2      # pairs None: the engine decides by batch size
        self.kstep, self.krows, self.kwaves = kstep, krows, kwaves    # K steps per pass (0: off, None: auto)
        self.num_steps = int(self.lengths.max())
        self.current_step = 0
        self.engine = None

    @property
    def cells(self):
        return sum(s.nx * s.ny for s in self.sims)

    def initialize(self):
        for s in self.sims:
            s.initialize_simulation(save=False, live=False)
            if s.grid_im is not None:
                raise NotImplementedError("batch members with complex sources (an imaginary grid): run them "
                                          "alone, or build them with complex_fields=False")
        self.engine = Fused2D([s.grid for s in self.sims], seg=self.seg, graph=self.graph, max_steps=self.num_steps,
                              pairs=self.pairs, seg2=self.seg2, kstep=self.kstep, krows=self.krows,
                              kwaves=self.kwaves)
        self.seg = self.engine.seg
        for s in self.sims:
            s.grid._f2d = False       # members advance through the batch engine only
        self.current_step = 0

    def reset(self):
        """Back to step 0 without re-planning: zero every member's fields and the monitor sums (series and
        FP64 DFT accumulators) in place. Materials, sources and monitor layouts are kept, so a new design can
        be written into the members' permittivity and run again at the cost of the steps alone."""
        if self.engine is None:
            self.initialize()
            return
        T = self.sims[0].backend.torch
        # every member's field buffers in one multi-tensor fill instead of six launches per member
        T._foreach_zero_([t for s in self.sims for fs in s.grid._fb if fs is not None for t in fs.values()])
        for s in self.sims:
            s.grid._cur = 0
            s.t = 0
            s.current_step = 0
        for buf in (self.engine.series, self.engine.dft):
            if buf is not None:
                buf.zero_()
        self.engine.set_active(None)
        self.current_step = 0

    def advance(self, count):
        """Run ``count`` steps of every member (bounded by the time axis)."""
        if self.engine is None:
            self.initialize()
        end = min(self.num_steps, self.current_step + int(count))
        first = True
        while self.current_step < end:
            n = self.current_step
            # members whose time axis has ended stay where they are; a run stops where the next one ends
            self.engine.set_active(self.lengths > n)
            stop = min(end, int(self.lengths[self.lengths > n].min()))
            c = min(self.chunk, stop - n)
            self.engine.run(n, c, coef=first)      # materials can only change between calls
            first = False
            self.current_step += c
        for s, ln in zip(self.sims, self.lengths):
            s.current_step = min(self.current_step, int(ln))
            s.t = s.current_step * s.dt

    def run(self, steps=None):
        """Initialise, run ``steps`` (default: the whole time axis) and finalise every member."""
        self.initialize()
        self.advance(self.num_steps if steps is None else steps)
        return [s.finalize_simulation() for s in self.sims]

    def synchronize(self):
        self.sims[0].synchronize()

    def autotune(self, steps=20, segs=None):
        """Pick the kernel shape with the lowest measured step time: (K, rows, waves) when the engine runs K
        steps per pass, else the tile height (the two-step kernel's tiles when it runs pairs, else the one-step
        kernel's); leaves the batch re-initialised."""
        T = self.sims[0].backend.torch
        if self.engine is None:
            self.initialize()
        if self.engine.kstep:
            best, timings = self.engine.tune_kstep(steps=steps, candidates=segs)
            self.kstep, self.krows, self.kwaves = best
            self.initialize()
            return best, timings
        pairs = self.engine.pairs
        steps += steps % 2
        best = None
        timings = {}
        for seg in segs or (self.engine.SEGS2 if pairs else self.SEGS):
            (self.engine.set_seg2 if pairs else self.engine.set_seg)(seg)
            self.engine.run(0, 2)                       # warm-up
            e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
            e0.record()
            self.engine.run(2, steps)
            e1.record()
            T.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            timings[seg] = ms
            if best is None or ms < timings[best]:
                best = seg
        if pairs:
            self.seg2 = best
        else:
            self.seg = best
        self.initialize()
        return best, timings
