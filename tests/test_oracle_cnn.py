"""CPU checks of the closure-network oracle: convolution orientation against a brute-force
loop, float64 autograd gradients against central finite differences."""
from __future__ import annotations

import numpy as np
import torch

from oracle import cnn3d as O


def test_conv_matches_brute_force_offsets():
    g = torch.Generator().manual_seed(1)
    x = torch.randn(3, 5, 4, 6, generator=g, dtype=torch.float64)
    w = torch.randn(2, 3, 3, 3, 3, generator=g, dtype=torch.float64)
    b = torch.randn(2, generator=g, dtype=torch.float64)
    flat = torch.cat([w.reshape(-1), b])
    y = O.forward(flat, x, (3, 2), ("none",), round_weights=False)
    assert float((y - O.brute_conv(x, w, b)).abs().max()) < 1e-12


def test_autograd_matches_finite_differences():
    g = torch.Generator().manual_seed(2)
    ch, acts = (2, 8, 1), ("relu", "softplus")
    n = sum(co * ci * 27 + co for ci, co in zip(ch[:-1], ch[1:]))
    flat = (0.3 * torch.randn(n, generator=g, dtype=torch.float64)).requires_grad_(True)
    x = torch.randn(2, 4, 4, 4, generator=g, dtype=torch.float64)
    dy = torch.randn(1, 4, 4, 4, generator=g, dtype=torch.float64)
    (O.forward(flat, x, ch, acts, round_weights=False) * dy).sum().backward()
    rng = np.random.default_rng(0)
    for i in rng.choice(n, 12, replace=False):
        e = torch.zeros(n, dtype=torch.float64)
        e[i] = 1e-6
        with torch.no_grad():
            fp = (O.forward(flat + e, x, ch, acts, round_weights=False) * dy).sum()
            fm = (O.forward(flat - e, x, ch, acts, round_weights=False) * dy).sum()
        fd = float(fp - fm) / 2e-6
        assert abs(fd - float(flat.grad[i])) < 1e-6 * max(1.0, abs(fd)), i


def test_weight_rounding_is_bf16_with_straight_through_gradient():
    w = torch.tensor([1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -9, -3.14159], dtype=torch.float64, requires_grad=True)
    r = O.bf16_round(w)
    # bf16 keeps 8 significant bits: ties to even
    assert r.detach().tolist() == [1.0, 1.0 + 2.0 ** -7, float(torch.tensor(-3.14159).to(torch.bfloat16))]
    (r * torch.tensor([0.3, -0.7, 1.9], dtype=torch.float64)).sum().backward()
    assert w.grad.tolist() == [0.3, -0.7, 1.9]
