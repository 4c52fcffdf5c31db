"""2D tracer particles on the MI355X (``aero2d_particles_step``): the pool after 25 steps of the
reference FlowSolver with a continuous rake (several releases per step, recycled rings) and a
pulsed line (``tests/golden/ns2d_particles.npz``, float64), graphs on and off."""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np
import pytest
import torch

from aerolume_amd.ns2d import solver as S
from aerolume_amd.ns2d.config import EmitterSpec
from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
EMITTERS = dict(
    rake=dict(position=(0.6, 1.2), kind="rake", end=(0.6, 2.8), count=5, species="particles", rate=400.0),
    line=dict(position=(1.2, 1.0), kind="line", end=(1.2, 3.0), count=6, species="particles", rate=1.0,
              pulse_period=0.15, pulse_duty=0.5, start=0.05))


def test_particles_match_reference():
    print("[particles gpu] pool vs reference", flush=True)
    for name in ("ns2d_particles.npz", "ns2d_pid_diag.npz"):
        if not (GOLDEN / name).exists():
            pytest.skip(f"missing fixture {name}")
    ref = dict(np.load(GOLDEN / "ns2d_particles.npz"))
    fx = dict(np.load(GOLDEN / "ns2d_pid_diag.npz"))
    pools = []
    for graphs in (True, False):
        cfg = cases.pid_diag_cfg("float64", device="cuda", emitters=[EmitterSpec(**EMITTERS["rake"]),
                                                                       EmitterSpec(**EMITTERS["line"])],
                                 particle_capacity=512)
        s = S.FlowSolver(dataclasses.replace(cfg, use_cuda_graphs=graphs))
        s.set_velocity(fx["u0"], fx["v0"])
        s.set_temperature(fx["theta0"])
        s.advance(25)
        assert s.t == pytest.approx(float(ref["t"]), rel=1e-12)
        p = s._particles
        assert np.array_equal(p.head.cpu().numpy(), ref["head"]), graphs
        assert np.array_equal(p.last.cpu().numpy(), ref["last"])
        assert np.allclose(p.acc.cpu().numpy(), ref["acc"], rtol=1e-10, atol=1e-13)
        alive = p.alive.cpu().numpy()
        assert np.array_equal(alive, ref["alive"])
        live = alive > 0
        assert live.sum() > 200
        assert np.abs(p.pos.cpu().numpy()[live] - ref["pos"][live]).max() < 1e-10
        assert np.abs(p.birth.cpu().numpy()[live] - ref["birth"][live]).max() < 1e-12
        snap = s.particles()
        assert np.array_equal(snap["nozzle"], ref["snap_nozzle"]) and np.array_equal(snap["emitter"], ref["snap_emitter"])
        assert len(s.particle_lines()) == int(ref["n_lines"])
        pools.append([t.clone() for t in p.buffers()])
    for a, b in zip(*pools):
        assert torch.equal(a, b)                 # captured and eager steps give the same pool
    s.reset()
    assert len(s.particles()["x"]) == 0


def test_particles_float32_close_to_float64():
    """The float32 pool after 25 steps against the float64 reference pool (same releases;
    positions to the float32 flow's accuracy)."""
    print("[particles gpu] float32", flush=True)
    if not (GOLDEN / "ns2d_particles.npz").exists():
        pytest.skip("missing fixture")
    ref = dict(np.load(GOLDEN / "ns2d_particles.npz"))
    fx = dict(np.load(GOLDEN / "ns2d_pid_diag.npz"))
    cfg = cases.pid_diag_cfg("float32", device="cuda", emitters=[EmitterSpec(**EMITTERS["rake"]),
                                                                   EmitterSpec(**EMITTERS["line"])],
                             particle_capacity=512, use_cuda_graphs=True)
    s = S.FlowSolver(cfg)
    s.set_velocity(fx["u0"], fx["v0"])
    s.set_temperature(fx["theta0"])
    s.advance(25)
    p = s._particles
    live = (p.alive.cpu().numpy() > 0) & (ref["alive"] > 0)
    assert live.sum() > 200
    assert np.abs(p.pos.cpu().double().numpy()[live] - ref["pos"][live]).max() < 2e-4
