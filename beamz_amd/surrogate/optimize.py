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
  with FDTD -> fine-tune on the verified samples; FDTD stays the judge of
  every accepted design.
"""
from __future__ import annotations

import time

import numpy as np


def train(net, rho, T, epochs=10, batch=256, lr=1e-3, seed=0, log=None):
    torch = net.torch
    rng = np.random.default_rng(seed)
    x = torch.as_tensor(rho, dtype=torch.float32, device=net.dev)
    y = torch.as_tensor(T, dtype=torch.float32, device=net.dev)
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


def relative_error(net, rho, T):
    torch = net.torch
    p = net.predict(torch.as_tensor(rho, dtype=torch.float32, device=net.dev)).cpu().numpy()
    return float(np.linalg.norm(p - T) / np.linalg.norm(T)), p


def screen(net, problem, rho_candidates, top=16, seed=0):
    """Verified mean transmission of the surrogate's top picks and of a random pick of the same size."""
    torch = net.torch
    x = torch.as_tensor(rho_candidates, dtype=torch.float32, device=net.dev)
    score = net.predict(x).mean(dim=1).cpu().numpy()
    best = np.argsort(-score)[:top]
    rand = np.random.default_rng(seed).choice(len(rho_candidates), top, replace=False)
    fom_best = problem.transmission(rho_candidates[best]).mean(axis=1)
    fom_rand = problem.transmission(rho_candidates[rand]).mean(axis=1)
    return fom_best, fom_rand, score[best]


class SurrogateLoop:
    """Surrogate-assisted optimisation of the broadband figure of merit (mean transmission)."""

    def __init__(self, net, verifier, candidates=1024, top=16, climb_steps=5, climb_rate=0.05, seed=0):
        self.net, self.verifier = net, verifier
        self.candidates, self.top = candidates, top
        self.climb_steps, self.climb_rate = climb_steps, climb_rate
        self.rng = np.random.default_rng(seed)
        self.pool_rho, self.pool_fom, self.pool_T = None, None, None
        self.history = []

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
        torch = self.net.torch
        score = self.net.predict(torch.as_tensor(cand, device=self.net.dev)).mean(dim=1).cpu().numpy()
        pick = cand[np.argsort(-score)[:self.top]]
        T = self.verifier.transmission(pick).astype(np.float32)
        fom = T.mean(axis=1)
        self.pool_rho = np.concatenate([self.pool_rho, pick])
        self.pool_T = np.concatenate([self.pool_T, T])
        self.pool_fom = np.concatenate([self.pool_fom, fom])
        keep = np.argsort(-self.pool_fom)[:256]
        self.pool_rho, self.pool_T, self.pool_fom = self.pool_rho[keep], self.pool_T[keep], self.pool_fom[keep]
        if len(self.pool_rho) >= 64:
            train(self.net, self.pool_rho, self.pool_T, epochs=finetune_epochs, batch=64, lr=lr)
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
