"""The closure network: 3x3x3 convolutions on a periodic grid, BF16 weights, FP32 accumulation,
forward / backward / Adam as HIP kernels (``aero_cnn_*`` of ``libaerolume_amd.so``)."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .._libcnn import CnnDesc

ACTS = {"none": 0, "relu": 1, "softplus": 2}
DEFAULT_CHANNELS = (9, 32, 32, 32, 32, 1)
DEFAULT_ACTS = ("relu", "relu", "relu", "relu", "softplus")


class CNN3D:
    """A ``aero_cnn`` plan for one periodic grid (nz, ny, nx). Parameters live on the device as a
    flat float32 vector: per layer w[co][ci][3][3][3] then b[co]."""

    def __init__(self, grid, channels=DEFAULT_CHANNELS, acts=DEFAULT_ACTS, device="cuda", seed: int = 0):
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the closure network runs on the GPU (HIP)")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        nz, ny, nx = (int(g) for g in grid)
        self.grid = (nz, ny, nx)
        self.channels = tuple(int(c) for c in channels)
        self.acts = tuple(acts)
        d = CnnDesc()
        d.nz, d.ny, d.nx = nz, ny, nx
        d.n_layers = len(self.acts)
        if len(self.channels) != d.n_layers + 1:
            raise ValueError("channels needs one entry more than acts")
        for i, c in enumerate(self.channels):
            d.channels[i] = c
        for i, a in enumerate(self.acts):
            d.act[i] = ACTS[a]
        h = ctypes.c_void_p()
        _lib.check(self.lib.aero_cnn_create(ctypes.byref(d), ctypes.byref(h)), "aero_cnn_create")
        self.handle = h
        self.nparam = int(self.lib.aero_cnn_param_count(h))
        self.set_params(self.init_params(seed))
        self.t = 0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.aero_cnn_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self.handle = None

    def _st(self):
        return ctypes.c_void_p(_lib.stream_handle(self.device))

    def layer_shapes(self):
        return [(co, ci) for ci, co in zip(self.channels[:-1], self.channels[1:])]

    def init_params(self, seed: int = 0) -> torch.Tensor:
        """He-normal weights (fan-in 27 ci), zero biases, seeded."""
        rng = np.random.default_rng(seed)
        parts = []
        for co, ci in self.layer_shapes():
            parts.append(rng.standard_normal(co * ci * 27) * np.sqrt(2.0 / (27 * ci)))
            parts.append(np.zeros(co))
        return torch.as_tensor(np.concatenate(parts), dtype=torch.float32, device=self.device)

    def split(self, flat: torch.Tensor):
        """[(w (co, ci, 3, 3, 3), b (co,)), ...] views of a flat parameter vector."""
        out, off = [], 0
        for co, ci in self.layer_shapes():
            w = flat[off:off + co * ci * 27].view(co, ci, 3, 3, 3)
            off += co * ci * 27
            b = flat[off:off + co]
            off += co
            out.append((w, b))
        return out

    def set_params(self, flat: torch.Tensor) -> None:
        flat = flat.to(device=self.device, dtype=torch.float32).contiguous()
        if flat.numel() != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {flat.numel()}")
        _lib.check(self.lib.aero_cnn_set_params(self.handle, self._st(), flat.data_ptr()), "aero_cnn_set_params")
        self.t = 0

    def params(self) -> torch.Tensor:
        out = torch.empty(self.nparam, device=self.device, dtype=torch.float32)
        _lib.check(self.lib.aero_cnn_get_params(self.handle, self._st(), out.data_ptr()), "aero_cnn_get_params")
        return out

    def grads(self) -> torch.Tensor:
        out = torch.empty(self.nparam, device=self.device, dtype=torch.float32)
        _lib.check(self.lib.aero_cnn_get_grads(self.handle, self._st(), out.data_ptr()), "aero_cnn_get_grads")
        return out

    def _check(self, x: torch.Tensor, c: int) -> torch.Tensor:
        if tuple(x.shape) != (c, *self.grid) or x.dtype != torch.float32 or x.device != self.device:
            raise ValueError(f"expected a float32 ({c}, {self.grid}) tensor on {self.device}")
        return x.contiguous()

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """Training forward (keeps the activations for :meth:`backward`)."""
        x = self._check(x, self.channels[0])
        y = torch.empty(self.channels[-1], *self.grid, device=self.device, dtype=torch.float32)
        _lib.check(self.lib.aero_cnn_forward(self.handle, self._st(), x.data_ptr(), y.data_ptr()), "aero_cnn_forward")
        return y

    def infer(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        x = self._check(x, self.channels[0])
        y = out if out is not None else torch.empty(self.channels[-1], *self.grid, device=self.device,
                                                     dtype=torch.float32)
        _lib.check(self.lib.aero_cnn_infer(self.handle, self._st(), x.data_ptr(), y.data_ptr()), "aero_cnn_infer")
        return y

    def backward(self, dy: torch.Tensor) -> None:
        """Parameter gradients of sum(dy * y) for the last :meth:`forward` (read with :meth:`grads`)."""
        dy = self._check(dy, self.channels[-1])
        _lib.check(self.lib.aero_cnn_backward(self.handle, self._st(), dy.data_ptr(), None), "aero_cnn_backward")

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        _lib.check(self.lib.aero_cnn_adam_step(self.handle, self._st(), lr, beta1, beta2, eps, self.t),
                   "aero_cnn_adam_step")
