"""CNN surrogate of the batched FDTD (new; no reference counterpart).

A 4-level encoder of 3x3 convolutions (32, 64, 128, 256 channels, two per
level, ReLU, 2x2 average pooling between levels) with a global-average-pool +
fully connected head maps a density map (B, H, W) in [0, 1] to the port
transmission at 8 wavelengths; back-propagating a weighting of the outputs
gives d(weighted transmission)/d(density). Every layer runs on the HIP kernels
of ``csrc/surrogate.hip``: BF16 weights and activations (NHWC), FP32
accumulation on the MFMA units, FP32 master weights and Adam state.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native as N

LEVELS = ((1, 32), (32, 64), (64, 128), (128, 256))


class Surrogate:
    def __init__(self, nout=8, seed=0, device="cuda"):
        import torch
        self.torch, self.dev, self.nout = torch, torch.device(device), int(nout)
        rng = np.random.default_rng(seed)
        P = {}
        for lvl, (ci, co) in enumerate(LEVELS):
            for s, (a, o) in enumerate(((ci, co), (co, co))):
                P[f"w{lvl}{s}"] = rng.standard_normal((o, 3, 3, a)) * np.sqrt(2.0 / (9 * a))   # He normal
                P[f"b{lvl}{s}"] = np.zeros(o)
        P["fw"] = rng.standard_normal((self.nout, LEVELS[-1][1])) * np.sqrt(1.0 / LEVELS[-1][1])
        P["fb"] = np.zeros(self.nout)
        # parameters, gradients and Adam moments each live in one flat buffer (every tensor a 256-byte aligned
        # view), so that zeroing the gradients and the Adam step are one launch each
        offs, o = {}, 0
        for k, v in P.items():
            offs[k] = o
            o += -(-v.size // 64) * 64
        self._flat = {n: torch.zeros(o, dtype=torch.float32, device=self.dev) for n in ("p", "g", "m", "v")}

        def views(buf):
            return {k: buf[offs[k]:offs[k] + v.size].view(v.shape) for k, v in P.items()}
        self.params, self.grads = views(self._flat["p"]), views(self._flat["g"])
        self.m, self.v = views(self._flat["m"]), views(self._flat["v"])
        for k, v in P.items():
            self.params[k].copy_(torch.from_numpy(np.asarray(v, dtype=np.float32)))
        # outputs are head outputs times out_scale (set from the training targets, so the head works at O(1))
        self.out_scale = torch.ones(self.nout, dtype=torch.float32, device=self.dev)
        self.step_count = 0
        self.work, self.work_for = None, None      # workspace of the ordered gradient reductions
        self.w16, self.wt16 = {}, {}
        for lvl, (ci, co) in enumerate(LEVELS):
            for s, a in enumerate((ci, co)):
                if lvl == 0 and s == 0:
                    continue
                k = f"w{lvl}{s}"
                self.w16[k] = torch.empty(self.params[k].shape, dtype=torch.int16, device=self.dev)
                self.wt16[k] = torch.empty((a, 3, 3, self.params[k].shape[0]), dtype=torch.int16, device=self.dev)
        self.sync_weights()

    # ---------------- helpers ----------------
    @property
    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def sync_weights(self):
        """bf16 copies (and channel-transposed copies for backward-data) of the FP32 conv weights, every layer
        in one launch (the pointer and size arrays are built once)."""
        if getattr(self, "_wcopy", None) is None:
            ks = list(self.w16)
            n = len(ks)
            P = ctypes.c_void_p * n
            I = ctypes.c_int * n
            self._wcopy = (P(*[self.params[k].data_ptr() for k in ks]), P(*[self.w16[k].data_ptr() for k in ks]),
                           P(*[self.wt16[k].data_ptr() for k in ks]), I(*[self.params[k].shape[0] for k in ks]),
                           I(*[self.params[k].shape[3] for k in ks]), n)
        w, w16, wt16, co, ci, n = self._wcopy
        N.call("bz_nn_weights_n", ctypes.addressof(w), ctypes.addressof(w16), ctypes.addressof(wt16),
               ctypes.addressof(co), ctypes.addressof(ci), n, self.stream)

    def load_state(self, params):
        for k, v in params.items():
            self.params[k].copy_(v)
        self.sync_weights()

    def _workspace(self, B, H, W):
        """Device FP32 workspace large enough for every layer's partial sums at this input size, so that the
        weight and bias gradients are added in a fixed order (deterministic training)."""
        if self.work_for != (B, H, W):
            L = N.lib()
            need = L.bz_nn_first_work(B, H, W, LEVELS[0][1])
            for lvl, (ci, co) in enumerate(LEVELS):
                h, w = H >> lvl, W >> lvl
                for a in ((ci, co) if lvl > 0 else (co,)):
                    need = max(need, L.bz_nn_wgrad_work(B, h, w, a, co, self._splits(B * h * w, co, a)))
            self.work = self.torch.empty(max(int(need), 1), dtype=self.torch.float32, device=self.dev)
            self.work_for = (B, H, W)
        return self.work

    def _empty16(self, *shape):
        return self.torch.empty(shape, dtype=self.torch.int16, device=self.dev)

    # ---------------- forward ----------------
    def forward(self, x):
        """x: (B, H, W) density (float tensor on the device; H, W divisible by 8). Returns out (B, nout) FP32."""
        T = self.torch
        x = x.to(device=self.dev, dtype=T.float32).contiguous()
        B, H, W = x.shape
        if H % 8 or W % 8:
            raise ValueError("input height and width must be multiples of 8")
        st = self.stream
        x16 = self._empty16(B, H, W, 1)
        N.call("bz_nn_to_bf16", x.data_ptr(), x16.data_ptr(), x.numel(), st)
        cache = {"x16": x16, "shape": (B, H, W), "acts": []}
        a, h, w = x16, H, W
        for lvl, (ci, co) in enumerate(LEVELS):
            if lvl > 0:
                p = self._empty16(B, h // 2, w // 2, ci)
                N.call("bz_nn_avgpool2", a.data_ptr(), p.data_ptr(), B, h, w, ci, st)
                a, h, w = p, h // 2, w // 2
            y0 = self._empty16(B, h, w, co)
            if lvl == 0:
                N.call("bz_nn_conv_first", a.data_ptr(), self.params["w00"].data_ptr(), self.params["b00"].data_ptr(),
                       y0.data_ptr(), B, h, w, co, st)
            else:
                N.call("bz_nn_conv3x3", a.data_ptr(), self.w16[f"w{lvl}0"].data_ptr(),
                       self.params[f"b{lvl}0"].data_ptr(), y0.data_ptr(), B, h, w, ci, co, 1, st)
            y1 = self._empty16(B, h, w, co)
            N.call("bz_nn_conv3x3", y0.data_ptr(), self.w16[f"w{lvl}1"].data_ptr(), self.params[f"b{lvl}1"].data_ptr(),
                   y1.data_ptr(), B, h, w, co, co, 1, st)
            cache["acts"].append((a, y0, y1, h, w))
            a = y1
        C = LEVELS[-1][1]
        gap = T.empty((B, C), dtype=T.float32, device=self.dev)
        out = T.empty((B, self.nout), dtype=T.float32, device=self.dev)
        N.call("bz_nn_head", a.data_ptr(), self.params["fw"].data_ptr(), self.params["fb"].data_ptr(), gap.data_ptr(),
               out.data_ptr(), B, h * w, C, self.nout, st)
        cache["gap"] = gap
        self.cache = cache
        return out * self.out_scale

    # ---------------- backward ----------------
    def backward(self, gout, input_grad=False):
        """Accumulates d(sum(out * gout))/d(param) into self.grads (zeroed first); returns the input gradient
        (B, H, W) FP32 if input_grad."""
        T = self.torch
        cache = self.cache
        st = self.stream
        self._flat["g"].zero_()
        B, H, W = cache["shape"]
        work = self._workspace(B, H, W)
        wp, wn = work.data_ptr(), work.numel()
        gout = (gout.to(device=self.dev, dtype=T.float32) * self.out_scale).contiguous()
        _, _, y1, h, w = cache["acts"][-1]
        C = LEVELS[-1][1]
        ga = self._empty16(B, h, w, C)
        N.call("bz_nn_head_bwd", gout.data_ptr(), cache["gap"].data_ptr(), self.params["fw"].data_ptr(),
               self.grads["fw"].data_ptr(), self.grads["fb"].data_ptr(), ga.data_ptr(), B, h * w, C, self.nout, st)
        gx = None
        for lvl in range(len(LEVELS) - 1, -1, -1):
            ci, co = LEVELS[lvl]
            a, y0, y1, h, w = cache["acts"][lvl]
            # the gradient at y1 through its ReLU: fused into the pooling backward below the top level
            if lvl < len(LEVELS) - 1:
                up = self._empty16(B, h, w, co)
                N.call("bz_nn_avgpool2_bwd_relu", ga.data_ptr(), y1.data_ptr(), up.data_ptr(), B, h, w, co, st)
                ga = up
            else:
                N.call("bz_nn_relu_mask", ga.data_ptr(), y1.data_ptr(), ga.numel(), st)
            splits = self._splits(B * h * w, co, co)
            # second convolution of the level: input y0 (co channels)
            N.call("bz_nn_conv3x3_wgrad", y0.data_ptr(), ga.data_ptr(), self.grads[f"w{lvl}1"].data_ptr(),
                   self.grads[f"b{lvl}1"].data_ptr(), B, h, w, co, co, splits, wp, wn, st)
            # its backward-data, through the ReLU of y0 (fused)
            g0 = self._empty16(B, h, w, co)
            N.call("bz_nn_conv3x3_dgrad_relu", ga.data_ptr(), self.wt16[f"w{lvl}1"].data_ptr(), y0.data_ptr(),
                   g0.data_ptr(), B, h, w, co, co, st)
            # first convolution of the level: input a (ci channels)
            if lvl == 0:
                if input_grad:
                    gx = T.empty((B, H, W), dtype=T.float32, device=self.dev)
                N.call("bz_nn_conv_first_bwd", a.data_ptr(), g0.data_ptr(), self.params["w00"].data_ptr(),
                       self.grads["w00"].data_ptr(), self.grads["b00"].data_ptr(),
                       gx.data_ptr() if gx is not None else None, B, h, w, co, wp, wn, st)
            else:
                N.call("bz_nn_conv3x3_wgrad", a.data_ptr(), g0.data_ptr(), self.grads[f"w{lvl}0"].data_ptr(),
                       self.grads[f"b{lvl}0"].data_ptr(), B, h, w, ci, co, self._splits(B * h * w, co, ci), wp, wn,
                       st)
                gi = self._empty16(B, h, w, ci)
                N.call("bz_nn_conv3x3_dgrad", g0.data_ptr(), self.wt16[f"w{lvl}0"].data_ptr(), gi.data_ptr(), B, h,
                       w, ci, co, st)
                ga = gi
        return gx

    @staticmethod
    def _splits(P, co, ci, target=2048):
        tiles = max(1, -(-co // 64)) * max(1, -(-(9 * ci) // 64))
        return int(max(1, min(-(-target // tiles), P // 256)))

    # ---------------- training / inference ----------------
    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        f = self._flat          # one elementwise update over every parameter (the padding stays zero)
        N.call("bz_nn_adam", f["p"].data_ptr(), f["g"].data_ptr(), f["m"].data_ptr(), f["v"].data_ptr(),
               f["p"].numel(), lr, b1, b2, eps, self.step_count, self.stream)
        self.sync_weights()

    def train_step(self, x, target, lr=1e-3):
        """One Adam step on the mean squared error of the outputs divided by out_scale; returns that loss
        (a device scalar)."""
        out = self.forward(x)
        t = target.to(device=self.dev, dtype=self.torch.float32)
        diff = (out - t) / self.out_scale
        self.backward(2.0 * diff / self.out_scale / diff.numel())
        self.adam(lr)
        return (diff * diff).mean()

    def predict(self, x, batch=256):
        outs = [self.forward(x[q:q + batch]) for q in range(0, x.shape[0], batch)]
        return self.torch.cat(outs)

    def input_gradient(self, x, weights):
        """d(sum_j weights[j] out[b, j]) / d x[b] for every sample b, (B, H, W)."""
        out = self.forward(x)
        w = self.torch.as_tensor(weights, dtype=self.torch.float32, device=self.dev)
        g = self.backward(w.expand_as(out).contiguous(), input_grad=True)
        return out, g
