"""HIP backend: the reference's Backend plugin on MI355X device memory.

Arrays are FP32 torch tensors on the GPU (torch is used only for device
allocation and streams). ``update_h_fields`` / ``update_e_field`` call the
hand-written kernels through the C ABI (bz_update_h_2d / bz_update_e_2d) and
update in place; the reference's NumPy backend returns new H arrays, and the
FDTD driver rebinds whatever is returned, so both contracts are met.

Complex dtypes raise ``TypeError`` in ``zeros``: the reference driver catches
exactly that and falls back to real fields (beamz/simulation/fdtd.py:165-183).
"""
from __future__ import annotations

import numpy as np

from ... import _native as N
from .base import Backend


def _torch():
    import torch
    return torch


class HipBackend(Backend):
    device = "hip"

    def __init__(self, device_index=0, **kwargs):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("beamz_amd HIP backend needs a ROCm GPU (torch.cuda.is_available() is False)")
        N.lib()
        self.torch = torch
        self.dev = torch.device("cuda", device_index)

    @property
    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def _check_dtype(self, dtype):
        if dtype is not None and np.dtype(dtype).kind == "c":
            raise TypeError("beamz_amd fields are real FP32; complex dtypes are not supported")

    def zeros(self, shape, dtype=None):
        self._check_dtype(dtype)
        return self.torch.zeros(shape, dtype=self.torch.float32, device=self.dev)

    def ones(self, shape, dtype=None):
        self._check_dtype(dtype)
        return self.torch.ones(shape, dtype=self.torch.float32, device=self.dev)

    def copy(self, array):
        return array.clone()

    def to_numpy(self, array):
        if isinstance(array, np.ndarray):
            return array
        return array.detach().cpu().numpy()

    def from_numpy(self, array):
        return self.torch.as_tensor(np.asarray(array, dtype=np.float32), device=self.dev).contiguous()

    def roll(self, array, shift, axis=None):
        return self.torch.roll(array, shift, dims=axis)

    @staticmethod
    def _pitch(t, name):
        if t.dim() != 2 or t.stride(1) != 1 or t.dtype != _torch().float32 or not t.is_cuda:
            raise ValueError(f"{name} must be a 2D FP32 device tensor with unit column stride")
        return t.stride(0)

    def update_h_fields(self, Hx, Hy, Ez, sigma, dx, dy, dt, mu_0, eps_0):
        ny, nx = Ez.shape
        if tuple(Hx.shape) != (ny, nx - 1) or tuple(Hy.shape) != (ny - 1, nx) or tuple(sigma.shape) != (ny, nx):
            raise ValueError("update_h_fields: shapes must be Hx (ny,nx-1), Hy (ny-1,nx), Ez/sigma (ny,nx)")
        N.call("bz_update_h_2d", Hx.data_ptr(), self._pitch(Hx, "Hx"), Hy.data_ptr(), self._pitch(Hy, "Hy"),
               Ez.data_ptr(), self._pitch(Ez, "Ez"), sigma.data_ptr(), self._pitch(sigma, "sigma"), nx, ny,
               float(dx), float(dy), float(dt), float(mu_0), float(eps_0), self.stream)
        return Hx, Hy

    def update_e_field(self, Ez, Hx, Hy, sigma, epsilon_r, dx, dy, dt, eps_0):
        ny, nx = Ez.shape
        if tuple(Hx.shape) != (ny, nx - 1) or tuple(Hy.shape) != (ny - 1, nx) or tuple(epsilon_r.shape) != (ny, nx):
            raise ValueError("update_e_field: shapes must be Hx (ny,nx-1), Hy (ny-1,nx), Ez/sigma/eps (ny,nx)")
        N.call("bz_update_e_2d", Ez.data_ptr(), self._pitch(Ez, "Ez"), Hx.data_ptr(), self._pitch(Hx, "Hx"),
               Hy.data_ptr(), self._pitch(Hy, "Hy"), sigma.data_ptr(), self._pitch(sigma, "sigma"),
               epsilon_r.data_ptr(), self._pitch(epsilon_r, "epsilon_r"), nx, ny, float(dx), float(dy), float(dt),
               float(eps_0), self.stream)
        return Ez
