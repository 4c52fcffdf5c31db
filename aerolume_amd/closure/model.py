"""The learned closure as the 3D solver sees it: before each step, features of the step-start
state (the cell-centred velocity gradient, scaled) go through the network, whose output
(scaled) is the eddy viscosity the stage kernels use where it lies in [0, trust]; elsewhere
the configured model (WALE) is used."""
from __future__ import annotations

import torch

from .network import CNN3D


class LearnedClosure:
    """``LearnedClosure(net, g_scale, nu_scale, trust)``: inputs g / g_scale, output
    nu_t = nu_scale * net(...), trusted in [0, trust]."""

    def __init__(self, net: CNN3D, g_scale: float, nu_scale: float, trust: float):
        self.net = net
        self.g_scale, self.nu_scale, self.trust = float(g_scale), float(nu_scale), float(trust)
        self.nut = None

    def bind(self, ops) -> None:
        if ops.dtype != torch.float32:
            raise ValueError("the learned closure runs with float32 solves")
        grid = (ops.nz, ops.ny, ops.nx)
        if grid != self.net.grid:
            raise ValueError(f"network grid {self.net.grid} does not match the solver's {grid}")
        self.feat = torch.empty(9, *grid, device=ops.device, dtype=torch.float32)
        self.out = torch.empty(1, *grid, device=ops.device, dtype=torch.float32)
        self.nut = torch.zeros(*grid, device=ops.device, dtype=torch.float32)

    def update(self, plan, state) -> None:
        plan.cell_gradient(state.u, state.v, state.w, self.feat, 1.0 / self.g_scale)
        self.net.infer(self.feat, self.out)
        torch.mul(self.out[0], self.nu_scale, out=self.nut)
