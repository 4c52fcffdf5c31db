"""CPU oracle of the learned-closure network (TEST INFRASTRUCTURE, never the product).

The network has no reference counterpart; this restates its definition with library
routines: 3x3x3 cross-correlations with circular padding (``torch.nn.functional.conv3d`` on a
circularly padded input), bias, ReLU / softplus, weights rounded to bfloat16 (round to nearest
even, ``Tensor.to(torch.bfloat16)``) and evaluated in float64. Gradients come from torch
autograd on the same graph.

Pinning: a brute-force loop over the 27 offsets on a tiny grid (fixes the offset orientation
w[co][ci][kz][ky][kx] <-> x[p + (kz-1, ky-1, kx-1)]), and central finite differences of the
float64 model against its autograd gradients.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def bf16_round(t: torch.Tensor) -> torch.Tensor:
    """Round to bfloat16 values in the forward pass; the gradient passes through unchanged
    (autograd through the cast itself would round the gradient to bfloat16 too)."""
    r = t.detach().to(torch.bfloat16).to(t.dtype)
    return t + (r - t.detach())


def split(flat: torch.Tensor, channels):
    out, off = [], 0
    for ci, co in zip(channels[:-1], channels[1:]):
        w = flat[off:off + co * ci * 27].view(co, ci, 3, 3, 3)
        off += co * ci * 27
        b = flat[off:off + co]
        off += co
        out.append((w, b))
    return out


def act(name: str, x: torch.Tensor) -> torch.Tensor:
    if name == "relu":
        return torch.relu(x)
    if name == "softplus":
        return torch.where(x > 20.0, x, torch.log1p(torch.exp(torch.clamp(x, max=20.0))))
    return x


def forward(flat: torch.Tensor, x: torch.Tensor, channels, acts, round_weights: bool = True) -> torch.Tensor:
    """x (C0, nz, ny, nx) -> (Cn, nz, ny, nx); flat parameters as the device layout."""
    h = x[None]
    for (w, b), a in zip(split(flat, channels), acts):
        if round_weights:
            w = bf16_round(w)
        hp = F.pad(h, (1, 1, 1, 1, 1, 1), mode="circular")
        h = act(a, F.conv3d(hp, w, b))
    return h[0]


def brute_conv(x: torch.Tensor, w: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """y[co][p] = b[co] + sum_ci sum_k w[co][ci][kz][ky][kx] x[ci][p + (kz-1, ky-1, kx-1)]
    with periodic wrap, by explicit rolls (small grids only)."""
    co_n, ci_n = w.shape[:2]
    y = b.view(-1, 1, 1, 1).expand(co_n, *x.shape[1:]).clone()
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                xs = torch.roll(x, shifts=(1 - kz, 1 - ky, 1 - kx), dims=(1, 2, 3))   # xs[p] = x[p + d]
                y = y + torch.einsum("oc,czyx->ozyx", w[:, :, kz, ky, kx], xs)
    return y
