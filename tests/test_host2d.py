"""Host-side helpers of the 2D solver on the CPU (no GPU needed): the probe interpolation and
the arc-length resampling behind ``FlowSolver.surface``, checked against closed forms."""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import torch

from aerolume_amd.ns2d.solver import bilinear, resample_closed


def test_bilinear_is_exact_on_linear_fields_and_clamps():
    ops = SimpleNamespace(x0=-1.0, y0=0.5, h=0.25)
    ny, nx = 6, 9
    for xoff, yoff in ((0.5, 0.5), (0.0, 0.5), (0.5, 0.0)):
        j, i = torch.meshgrid(torch.arange(ny, dtype=torch.float64), torch.arange(nx, dtype=torch.float64),
                              indexing="ij")
        X, Y = ops.x0 + (i + xoff) * ops.h, ops.y0 + (j + yoff) * ops.h
        f = 3.0 * X - 2.0 * Y + 0.7
        g = torch.Generator().manual_seed(3)
        x = ops.x0 + (xoff + (nx - 1) * torch.rand(50, generator=g, dtype=torch.float64)) * ops.h
        y = ops.y0 + (yoff + (ny - 1) * torch.rand(50, generator=g, dtype=torch.float64)) * ops.h
        assert torch.allclose(bilinear(f, x, y, ops, xoff, yoff), 3.0 * x - 2.0 * y + 0.7, rtol=0, atol=1e-13)
        # grid points reproduce the samples exactly, points outside take the edge value
        assert float(bilinear(f, X[2:3, 4], Y[2:3, 4], ops, xoff, yoff)) == float(f[2, 4])
        far = bilinear(f, torch.tensor([-50.0]), torch.tensor([50.0]), ops, xoff, yoff)
        assert float(far) == float(f[ny - 1, 0])


def test_bilinear_product_term():
    """A bilinear field x*y is reproduced exactly inside a cell (the cross term is kept)."""
    ops = SimpleNamespace(x0=0.0, y0=0.0, h=1.0)
    j, i = torch.meshgrid(torch.arange(4, dtype=torch.float64), torch.arange(4, dtype=torch.float64), indexing="ij")
    f = (i + 0.5) * (j + 0.5)
    x, y = torch.tensor([1.3, 2.9]), torch.tensor([2.2, 0.6])
    assert torch.allclose(bilinear(f, x.double(), y.double(), ops), (x * y).double(), atol=1e-14)


def test_resample_closed_square_and_circle():
    sq = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    pts, arc = resample_closed(sq, 8)
    assert np.allclose(arc, np.arange(8.0))
    assert np.allclose(pts, [[0, 0], [1, 0], [2, 0], [2, 1], [2, 2], [1, 2], [0, 2], [0, 1]])
    th = np.linspace(0, 2 * math.pi, 400, endpoint=False)
    circ = np.c_[np.cos(th), np.sin(th)]
    pts, arc = resample_closed(circ, 37)
    seg = np.hypot(*np.diff(np.vstack([pts, pts[:1]]), axis=0).T)
    assert np.allclose(seg, seg.mean(), rtol=1e-3) and np.allclose(np.hypot(*pts.T), 1.0, atol=1e-4)
    assert arc[0] == 0.0 and np.all(np.diff(arc) > 0)


def test_particle_polylines_group_streaklines_and_timelines():
    """``TracerParticles.lines``: one streakline per continuous nozzle (youngest first) and one
    timeline per pulsed release (in nozzle order), from a hand-filled pool on the CPU."""
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d.config import EmitterSpec, SolverConfig
    from aerolume_amd.ns2d.particles import TracerParticles
    cfg = SolverConfig(nx=16, ny=8, h=0.25, x0=0.0, y0=0.0, re=100.0, dtype="float64", device="cpu", dye=False,
                       perturb=False)
    ops = su.make_operators(cfg, device="cpu")
    ems = [EmitterSpec(position=(0.5, 0.5), kind="rake", end=(0.5, 1.5), count=2, species="particles", rate=10.0),
           EmitterSpec(position=(1.0, 0.5), kind="line", end=(1.0, 1.5), count=3, species="particles",
                       pulse_period=0.5)]
    tp = TracerParticles(ems, ops, capacity=40)          # 5 nozzles x 8 slots
    assert tp.ring == 8 and tp.capacity == 40 and tp.pulsed.tolist() == [0.0, 1.0]
    t = 1.0
    # nozzle 0: three particles of ages 0.3, 0.1, 0.2; nozzle 1: one (no line)
    for slot, (x, age) in zip((0, 1, 2), ((0.8, 0.3), (0.6, 0.1), (0.7, 0.2))):
        tp.pos[slot] = torch.tensor([x, 0.5], dtype=torch.float64)
        tp.birth[slot], tp.alive[slot] = t - age, 1.0
    tp.pos[8], tp.birth[8], tp.alive[8] = torch.tensor([0.9, 1.5], dtype=torch.float64), t - 0.2, 1.0
    # pulsed nozzles 2..4: two releases (births 0.5 and 0.9), written in shuffled nozzle order
    for n, b, x in ((4, 0.5, 1.4), (2, 0.5, 1.2), (3, 0.5, 1.3), (3, 0.9, 1.05), (2, 0.9, 1.02)):
        slot = n * 8 + int(b * 10) % 8
        tp.pos[slot] = torch.tensor([x, 0.5 + 0.5 * (n - 2)], dtype=torch.float64)
        tp.birth[slot], tp.alive[slot] = b, 1.0
    lines = tp.lines(t)
    assert len(lines) == 3
    assert np.allclose(lines[0][:, 0], [0.6, 0.7, 0.8])                  # youngest first
    tl = sorted(lines[1:], key=len)
    assert np.allclose(tl[0][:, 1], [0.5, 1.0]) and np.allclose(tl[1][:, 1], [0.5, 1.0, 1.5])
    snap = tp.snapshot(t)
    assert len(snap["x"]) == 9 and set(snap["emitter"].tolist()) == {0, 1}
