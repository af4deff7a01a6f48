"""Training data and objective of the surrogate, produced by the batched GPU FDTD.

The device is the batched-2D member geometry at a fixed size: a 256 x 256 cell
SiO2 domain (25 nm cells, 10-cell PML), 0.5 um Si input and output guides and a
128 x 128 pixel (3.2 um) design region between them, a Gaussian-pulse point
source in the input guide and a DFT line monitor across the output guide at 8
wavelengths in 1.50-1.60 um, 2000 steps. The label of a density map rho is its
mode-overlap transmission spectrum

    T_f(rho) = |a_f(rho)|^2 / |a_f(straight guide)|^2,  a_f = sum_p c_fp Ez_hat(p, w_f)

with c_f the conjugated, normalised fundamental slab mode of the output guide
at wavelength f; the broadband figure of merit is mean_f T_f.
"""
from __future__ import annotations

import numpy as np

from ..const import µm
from ..workloads import EPS_OX, EPS_SI, batch_member_2d, binary_densities

PX = 128
SIDE = 256


class TransmissionProblem:
    """A fixed batch of `batch` simulations of the member geometry, re-used for every evaluation."""

    def __init__(self, batch=256, steps=2000, seed=0, side=SIDE, px=PX):
        import beamz_amd as bz
        from ..optimization.adjoint import mode_weights
        from ..simulation.batch import FDTDBatch
        self.px, self.side, self.steps = px, side, steps
        rng = np.random.default_rng(seed)
        self.sims, self.mons = [], []
        for _ in range(batch):
            mesh, t, (m_in, m_out) = batch_member_2d(side, rng, steps=steps, design_um=px * 25e-9 / µm)
            self.sims.append(bz.FDTD(mesh, time=t))
            self.mons.append(m_out)
        s0 = self.sims[0]
        self.dx = s0.dx
        self.x0 = self.y0 = (side - px) // 2
        self.lams = bz.LIGHT_SPEED / self.mons[0].dft_frequencies
        H = side * s0.dy
        self.c = np.stack([mode_weights(self.mons[0], s0.dx, s0.dy, H / 2, 0.5 * µm, 3.476, 1.444, lam)
                           for lam in self.lams])
        self.batch = FDTDBatch(self.sims)
        self.torch = s0.backend.torch
        self.batch.initialize()
        if len(self.sims) >= 64:
            self.batch.autotune(steps=20)     # one problem object serves every evaluation: tune the tiles once
        T = self.torch
        self.plans = [next(p for p in s._plans if p.monitor is m) for s, m in zip(self.sims, self.mons)]
        self.npts = self.plans[0].layout[1]
        self.c_dev = T.from_numpy(self.c).to(s0.backend.dev)
        self.ref = None
        straight = np.zeros((1, px, px), dtype=np.float32)
        band = (np.arange(px) + 0.5 - px / 2) * s0.dy
        straight[0, np.abs(band) < 0.25 * µm, :] = 1.0
        self.ref = self.raw_overlaps(self.torch.tensor(straight))[0]          # |a_f|^2 of the straight guide

    @property
    def box(self):
        return (self.x0, self.y0, self.px, self.px)

    def eps_of(self, rho):
        return EPS_OX + rho * (EPS_SI - EPS_OX)

    def submit(self, rho):
        """Start the evaluation of at most len(self.sims) density maps (m, px, px) in [0, 1]: reset, write the
        permittivities and enqueue every step. Returns at once; collect(m) waits and reads the result."""
        T = self.torch
        chunk = T.as_tensor(rho, dtype=T.float32).to(self.sims[0].backend.dev)
        if not 0 < chunk.shape[0] <= len(self.sims):
            raise ValueError(f"a chunk holds 1..{len(self.sims)} density maps")
        self.batch.reset()
        # permittivities of the whole chunk in one op, written into the members' design boxes in one
        # multi-tensor copy (members past the chunk repeat its first map)
        e = self.eps_of(chunk)
        src = [e[k] if k < chunk.shape[0] else e[0] for k in range(len(self.sims))]
        T._foreach_copy_([s.grid.eps[self.y0:self.y0 + self.px, self.x0:self.x0 + self.px] for s in self.sims], src)
        self.batch.advance(self.batch.num_steps)

    def collect(self, m):
        """|a_f|^2 (m, nf) of the first m maps of the last submitted chunk."""
        T = self.torch
        z = T.stack([p.dft[:, :self.npts, :] for p in self.plans[:m]])   # (m, nf, npts, 2)
        a = (self.c_dev * T.complex(z[..., 0], z[..., 1])).sum(dim=2)
        return (a.real ** 2 + a.imag ** 2).cpu().numpy()

    def raw_overlaps(self, rho):
        """|a_f|^2 for each density map of rho (n, px, px) in [0, 1] (torch or numpy), in batches."""
        B = len(self.sims)
        out = []
        for q0 in range(0, len(rho), B):
            self.submit(rho[q0:q0 + B])
            out.append(self.collect(min(B, len(rho) - q0)))
        return np.concatenate(out)

    def transmission(self, rho):
        return self.raw_overlaps(rho) / self.ref


def random_densities(n, seed, px=PX):
    """The batched workload's density generator: seeded Bernoulli(0.5), 3-pixel Gaussian blur, threshold."""
    return binary_densities(n, px, np.random.default_rng(seed)).astype(np.float32)


def generate(problem, n, seed=0):
    """n (density, transmission spectrum) pairs from the batched FDTD: the densities of random_densities(n,
    seed), each chunk's drawn on the host while the GPU runs the previous chunk."""
    rng = np.random.default_rng(seed)
    B = len(problem.sims)

    def draw(m):
        return binary_densities(m, problem.px, rng).astype(np.float32)

    rho, T = [], []
    cur = draw(min(B, n)) if n else None
    done = 0
    while cur is not None:
        problem.submit(cur)
        done += len(cur)
        nxt = draw(min(B, n - done)) if done < n else None
        T.append(problem.collect(len(cur)) / problem.ref)
        rho.append(cur)
        cur = nxt
    if not rho:
        return np.zeros((0, problem.px, problem.px), np.float32), np.zeros((0, len(problem.ref)), np.float32)
    return np.concatenate(rho), np.concatenate(T).astype(np.float32)
