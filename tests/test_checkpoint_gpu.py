"""Checkpoints in the reference's format (ns2d/solver.py:1280-1336, ns3d/solver.py:1738-1789):
a checkpoint written by the reference's own FlowSolver / FlowSolver3D (fixtures written by
``tests/golden/make_checkpoint_fixtures.py``) resumes here and continues as the reference
did; a checkpoint written here carries the reference's keys and types and restores a fresh
solver bit for bit."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _unflatten(F) -> dict:
    """The reference's state_dict from its flattened fixture."""
    ck, hist = {}, {}
    for k in F.files:
        if k.startswith("ckpt:history:"):
            hist[k.split(":", 2)[2]] = F[k]
        elif k.startswith("ckpt:"):
            v = F[k]
            ck[k[5:]] = v.item() if v.ndim == 0 and k[5:] in ("version", "steps", "re") else torch.as_tensor(v)
    ck["history"] = hist
    return ck


def _cfg2d():
    from aerolume_amd.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig
    return SolverConfig(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=150.0,
                        bodies=[ImmersedBody(cases.circle(64, 0.75, 2.0, 2.0), thermal="temperature",
                                             thermal_value=1.0)],
                        dtype="float64", device="cuda", dye=True, dye_inlet=False, perturb=True,
                        thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                        emitters=[EmitterSpec(position=(0.8, 1.8), channel=0, rate=1.0),
                                  EmitterSpec(position=(0.8, 2.2), channel=1, rate=0.5)])


def _cfg3d():
    from aerolume_amd.ns3d.config import SolverConfig3D
    return SolverConfig3D(nx=16, ny=12, nz=10, h=2 * math.pi / 16, x0=0.0, y0=0.0, z0=0.0, re=400.0, periodic=True,
                          turbulence="wale", dtype="float64", device="cuda", dye=False, perturb=False)


def _err(a, b):
    return float(np.abs(np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
                        - np.asarray(b, dtype=np.float64)).max())


def test_reference_checkpoint_resumes_2d(golden):
    """The reference's checkpoint after 12 steps (two smoke channels, heated disc) loads into
    FlowSolver; 5 more steps equal the reference's to 1e-10 and the history continues its."""
    _need_gpu()
    from aerolume_amd.ns2d.solver import FlowSolver
    F = golden("ckpt_ns2d.npz")
    ck = _unflatten(F)
    assert ck["version"] == 2 and ck["dye_extra"].shape[0] == 1 and ck["t"].dim() == 0
    fs = FlowSolver(_cfg2d())
    fs.load_state_dict(ck)
    assert fs.steps == 12 and len(fs.history()["t"]) == 12
    fs.advance(5)
    st = fs.state
    for k in ("u", "v", "p", "dye", "dye_extra", "theta", "t"):
        ref = F[f"after:{k}"]
        assert _err(getattr(st, k), ref) < 1e-10 * max(1.0, float(np.abs(ref).max())), k
    h = fs.history()
    for k in ("t", "cd", "cl", "cm"):
        assert _err(h[k], F[f"after:history:{k}"]) < 1e-10, k


def test_checkpoint_round_trip_2d():
    """state_dict carries the reference's keys (version 2, 0-d t / dt tensors, dye and
    dye_extra split, body_coefficients, history, re); loading it into a fresh solver restores
    the state bit for bit and the two continue identically."""
    _need_gpu()
    from aerolume_amd.ns2d.solver import FlowSolver
    a = FlowSolver(_cfg2d())
    a.advance(13)
    ck = a.state_dict()
    assert set(ck) == {"version", "steps", "re", "history", "body_coefficients", "u", "v", "p", "t", "dt", "dye",
                       "dye_extra", "theta"}
    assert ck["version"] == 2 and ck["t"].dim() == 0 and ck["dt"].dim() == 0 and ck["dye"].dim() == 2
    assert ck["dye_extra"].shape == (1, 32, 64) and ck["body_coefficients"].shape == (1, 3)
    assert all(not t.is_cuda for t in ck.values() if torch.is_tensor(t))
    b = FlowSolver(_cfg2d())
    b.load_state_dict(ck)
    for k in ("u", "v", "p", "t", "dt", "dye", "dye_extra", "theta"):
        assert torch.equal(getattr(a.state, k), getattr(b.state, k)), k
    a.advance(7)
    b.advance(7)
    torch.cuda.synchronize()
    for k in ("u", "v", "p", "dye", "dye_extra", "theta"):
        assert torch.equal(getattr(a.state, k), getattr(b.state, k)), k
    assert np.array_equal(a.history()["cd"], b.history()["cd"])


def test_reference_checkpoint_resumes_3d(golden):
    """The reference's FlowSolver3D checkpoint after 6 steps of a periodic WALE box loads into
    FlowSolver3D; 4 more steps equal the reference's to 1e-10."""
    _need_gpu()
    from aerolume_amd.ns3d.solver import FlowSolver3D
    F = golden("ckpt_ns3d.npz")
    ck = _unflatten(F)
    assert ck["version"] == 1 and ck["t"].dim() == 0
    fs = FlowSolver3D(_cfg3d())
    fs.load_state_dict(ck)
    assert fs.steps == 6 and len(fs.history()["t"]) == 6
    fs.advance(4)
    st = fs.state
    for k in ("u", "v", "w", "p", "t"):
        ref = F[f"after:{k}"]
        assert _err(getattr(st, k), ref) < 1e-10 * max(1.0, float(np.abs(ref).max())), k


def test_checkpoint_round_trip_3d():
    _need_gpu()
    from aerolume_amd.ns3d.solver import FlowSolver3D
    a = FlowSolver3D(_cfg3d())
    rng = np.random.default_rng(5)
    a.set_velocity(*(0.2 * rng.standard_normal(tuple(getattr(a.state, k).shape)) for k in ("u", "v", "w")))
    a.advance(5)
    ck = a.state_dict()
    assert {"version", "steps", "re", "history", "body_coefficients", "u", "v", "w", "p", "t", "dt"} <= set(ck)
    assert ck["version"] == 1 and ck["t"].dim() == 0 and len(ck["history"]["t"]) == 5
    b = FlowSolver3D(_cfg3d())
    b.load_state_dict(ck)
    a.advance(3)
    b.advance(3)
    torch.cuda.synchronize()
    for k in ("u", "v", "w", "p", "t"):
        assert torch.equal(getattr(a.state, k), getattr(b.state, k)), k
    assert len(b.history()["t"]) == 8
