"""Training, screening and the surrogate-assisted design loop.

* ``train``: Adam on the mean squared error of the transmission spectrum, on
  the GPU kernels of the surrogate.
* ``relative_error``: relative L2 error of the predicted spectra against FDTD.
* ``screen``: rank candidate density maps by the surrogate's broadband figure
  of merit (mean transmission) and verify the top ones with the batched FDTD;
  returns the verified figures of the surrogate's picks and of an equally sized
  random pick, which is what the surrogate gains over screening without it.
* ``SurrogateLoop``: rounds of propose (mutations of the best verified designs
  climbed along the surrogate's input gradient) -> rank -> verify the top-k
  with FDTD -> fine-tune on the verified samples; FDTD stays the arbiter of
  every accepted design.
"""
from __future__ import annotations

import time

import numpy as np


def train(net, rho, T, epochs=10, batch=256, lr=1e-3, seed=0, log=None, set_scale=True):
    """Mini-batch Adam on the spectra. With set_scale the network's output scale becomes the mean magnitude
    of the targets per wavelength (the loss is then a relative squared error, whatever their level)."""
    torch = net.torch
    rng = np.random.default_rng(seed)
    x = torch.as_tensor(rho, dtype=torch.float32, device=net.dev)
    y = torch.as_tensor(T, dtype=torch.float32, device=net.dev)
    if set_scale:
        net.out_scale = y.abs().mean(dim=0).clamp_min(1e-12).contiguous()
    n = x.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(net.dev)
        losses = []
        for q in range(0, n - batch + 1, batch):
            idx = perm[q:q + batch]
            losses.append(net.train_step(x[idx], y[idx], lr=lr))
        loss = float(torch.stack(losses).mean()) if losses else float("nan")
        hist.append(loss)
        if log:
            log(f"epoch {ep}: mse {loss:.4e}")
    return hist


LOG_FLOOR = 1e-4


def to_log(T):
    """Log-space targets: transmissions of random designs span orders of magnitude."""
    return np.log(np.asarray(T, dtype=np.float64) + LOG_FLOOR).astype(np.float32)


def from_log(y):
    return np.exp(np.asarray(y, dtype=np.float64)) - LOG_FLOOR


def predict_T(net, rho, log_space=False):
    torch = net.torch
    p = net.predict(torch.as_tensor(rho, dtype=torch.float32, device=net.dev)).cpu().numpy()
    return from_log(p) if log_space else p


def relative_error(net, rho, T, log_space=False):
    p = predict_T(net, rho, log_space)
    return float(np.linalg.norm(p - T) / np.linalg.norm(T)), p


def screen(net, problem, rho_candidates, top=16, seed=0, log_space=False):
    """Verified mean transmission of the surrogate's top picks and of a random pick of the same size."""
    score = predict_T(net, rho_candidates, log_space).mean(axis=1)
    best = np.argsort(-score)[:top]
    rand = np.random.default_rng(seed).choice(len(rho_candidates), top, replace=False)
    fom_best = problem.transmission(rho_candidates[best]).mean(axis=1)
    fom_rand = problem.transmission(rho_candidates[rand]).mean(axis=1)
    return fom_best, fom_rand, score[best]


class PlainAdjointLoop:
    """The method without the surrogate: batched exact-adjoint gradient ascent of the broadband figure of
    merit mean_f T_f over `designs` density maps (cone filter r=3, tanh projection beta=8, Adam on the
    latent densities), every iteration one batched forward + adjoint FDTD."""

    def __init__(self, problem, designs=16, lr=0.05, radius=3, beta=8.0, seed=0):
        import beamz_amd as bz
        from ..optimization.adjoint import AdjointProblem2D
        from ..workloads import batch_member_2d
        from ..const import µm
        torch = problem.torch
        rng = np.random.default_rng(seed)
        sims, mons = [], []
        for _ in range(designs):
            mesh, t, (m_in, m_out) = batch_member_2d(problem.side, rng, steps=problem.steps,
                                                     design_um=problem.px * 25e-9 / µm)
            sims.append(bz.FDTD(mesh, time=t))
            mons.append(m_out)
        fscale = 1.0 / (len(problem.ref) * problem.ref)
        self.adj = AdjointProblem2D(sims, mons, problem.c, problem.box, fscale=fscale)
        self.problem, self.radius, self.beta, self.lr = problem, radius, beta, lr
        self.rho = torch.tensor(np.clip(0.5 + 0.05 * rng.standard_normal((designs, problem.px, problem.px)), 0, 1),
                                dtype=torch.float32, device=problem.sims[0].backend.dev, requires_grad=True)
        self.opt = torch.optim.Adam([self.rho], lr=lr)
        self.history = []

    def physical(self):
        from ..optimization.adjoint import cone_filter_torch, tanh_projection_torch
        return tanh_projection_torch(cone_filter_torch(self.rho, self.radius), self.beta)

    def step(self):
        from ..optimization.adjoint import FDTDObjective
        torch = self.problem.torch
        self.opt.zero_grad()
        phys = self.physical()
        J = FDTDObjective.apply(self.problem.eps_of(phys), self.adj)
        (-J.sum()).backward()
        self.opt.step()
        with torch.no_grad():
            self.rho.clamp_(0.0, 1.0)
        return J.detach().cpu().numpy(), phys.detach()

    def run(self, iterations=40, clock=None):
        t0 = time.perf_counter() if clock is None else clock
        best, best_rho = -1.0, None
        for _ in range(iterations):
            J, phys = self.step()
            k = int(np.argmax(J))
            if J[k] > best:
                best, best_rho = float(J[k]), phys[k].cpu().numpy()
            self.history.append((time.perf_counter() - t0, best))
        return best, best_rho


class SurrogateLoop:
    """Surrogate-assisted optimisation of the broadband figure of merit (mean transmission)."""

    def __init__(self, net, verifier, candidates=1024, top=16, climb_steps=5, climb_rate=0.05, seed=0,
                 log_space=False):
        self.net, self.verifier, self.log_space = net, verifier, log_space
        self.candidates, self.top = candidates, top
        self.climb_steps, self.climb_rate = climb_steps, climb_rate
        self.rng = np.random.default_rng(seed)
        self.pool_rho, self.pool_fom, self.pool_T = None, None, None
        self.history = []
        self.rounds = []          # per round: the picks' verified and predicted spectra, compared

    def propose(self):
        torch = self.net.torch
        k = len(self.pool_rho)
        parents = self.pool_rho[self.rng.integers(0, k, self.candidates)]
        noise = self.rng.standard_normal(parents.shape).astype(np.float32) * 0.15
        x = torch.as_tensor(np.clip(parents + noise, 0.0, 1.0), device=self.net.dev)
        w = np.full(self.net.nout, 1.0 / self.net.nout, dtype=np.float32)
        for _ in range(self.climb_steps):                 # ascend the surrogate's broadband FOM
            for q in range(0, x.shape[0], 256):
                _, g = self.net.input_gradient(x[q:q + 256], w)
                gn = g / (g.abs().amax(dim=(1, 2), keepdim=True) + 1e-12)
                x[q:q + 256] = (x[q:q + 256] + self.climb_rate * gn).clamp(0.0, 1.0)
        return x.cpu().numpy()

    def step(self, finetune_epochs=3, lr=5e-4):
        cand = self.propose()
        pred = predict_T(self.net, cand, self.log_space)
        score = pred.mean(axis=1)
        order = np.argsort(-score)[:self.top]
        pick = cand[order]
        T = self.verifier.transmission(pick).astype(np.float32)
        fom = T.mean(axis=1)
        # the surrogate's error against FDTD on the designs it picked (before it is fine-tuned on them)
        self.rounds.append({"verified_fom_mean": float(fom.mean()), "verified_fom_max": float(fom.max()),
                            "predicted_fom_mean": float(score[order].mean()),
                            "rel_error": float(np.linalg.norm(pred[order] - T) / max(np.linalg.norm(T), 1e-30))})
        self.pool_rho = np.concatenate([self.pool_rho, pick])
        self.pool_T = np.concatenate([self.pool_T, T])
        self.pool_fom = np.concatenate([self.pool_fom, fom])
        keep = np.argsort(-self.pool_fom)[:256]
        self.pool_rho, self.pool_T, self.pool_fom = self.pool_rho[keep], self.pool_T[keep], self.pool_fom[keep]
        if len(self.pool_rho) >= 64:
            y = to_log(self.pool_T) if self.log_space else self.pool_T
            train(self.net, self.pool_rho, y, epochs=finetune_epochs, batch=64, lr=lr, set_scale=False)
        return float(fom.max())

    def run(self, seed_rho, seed_T, rounds=10, target=None, clock=None):
        t0 = time.perf_counter() if clock is None else clock
        self.pool_rho, self.pool_T = np.asarray(seed_rho, np.float32), np.asarray(seed_T, np.float32)
        self.pool_fom = self.pool_T.mean(axis=1)
        best = float(self.pool_fom.max())
        for r in range(rounds):
            best = max(best, self.step())
            self.history.append((time.perf_counter() - t0, best))
            if target is not None and best >= target:
                break
        return best
