"""Write checkpoint fixtures from the reference's own FlowSolver / FlowSolver3D state_dict()
(run in the build container only; imports the reference from ``/root/reference/src``,
read-only, on the CPU). Each checkpoint is flattened into an ``.npz`` (keys ``ckpt:<name>``,
``ckpt:history:<column>``; scalars as 0-d arrays) so it loads with numpy alone, together with
the reference's state after resuming from it for a few more steps.

* ``ckpt_ns2d.npz`` - 64x32 tunnel, disc held at theta = 1 (Pr 0.71, Ri 0.1), two smoke
  channels from point emitters (no inlet comb), float64: the checkpoint after 12 steps, then
  u, v, p, dye, dye_extra, theta, t after 5 more steps and the coefficient history.
* ``ckpt_ns3d.npz`` - 16x12x10 triply periodic WALE box from a seeded noisy field, float64:
  the checkpoint after 6 steps, then u, v, w, p, t after 4 more steps.

Usage: ``python tests/golden/make_checkpoint_fixtures.py``.
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent
sys.path.insert(0, str(REF))

import torch  # noqa: E402

from aerolume.solvers.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig  # noqa: E402
from aerolume.solvers.ns2d.solver import FlowSolver  # noqa: E402
from aerolume.solvers.ns3d import core as core3  # noqa: E402
from aerolume.solvers.ns3d.config import SolverConfig3D  # noqa: E402
from aerolume.solvers.ns3d.solver import FlowSolver3D  # noqa: E402


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def flatten(ckpt: dict) -> dict:
    out = {}
    for k, v in ckpt.items():
        if k == "history":
            for col, arr in v.items():
                out[f"ckpt:history:{col}"] = np.asarray(arr)
        elif torch.is_tensor(v):
            out[f"ckpt:{k}"] = v.numpy()
        else:
            out[f"ckpt:{k}"] = np.asarray(v)
    return out


def np_(t):
    return t.detach().cpu().numpy()


def fixture_2d():
    cfg = SolverConfig(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=150.0,
                       bodies=[ImmersedBody(circle(64, 0.75, 2.0, 2.0), thermal="temperature", thermal_value=1.0)],
                       dtype="float64", device="cpu", dye=True, dye_inlet=False, perturb=True,
                       thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                       emitters=[EmitterSpec(position=(0.8, 1.8), channel=0, rate=1.0),
                                 EmitterSpec(position=(0.8, 2.2), channel=1, rate=0.5)])
    fs = FlowSolver(cfg)
    fs.advance(12)
    arrs = flatten(fs.state_dict())
    fs.advance(5)
    st = fs.state
    arrs.update({"after:u": np_(st.u), "after:v": np_(st.v), "after:p": np_(st.p), "after:dye": np_(st.dye),
                 "after:dye_extra": np_(st.dye_extra), "after:theta": np_(st.theta), "after:t": np_(st.t)})
    for k, v in fs.history().items():
        arrs[f"after:history:{k}"] = np.asarray(v)
    np.savez_compressed(OUT / "ckpt_ns2d.npz", **arrs)


def fixture_3d():
    n = (16, 12, 10)
    cfg = SolverConfig3D(nx=n[0], ny=n[1], nz=n[2], h=2 * math.pi / 16, x0=0.0, y0=0.0, z0=0.0, re=400.0,
                         periodic=True, turbulence="wale", dtype="float64", device="cpu", dye=False, perturb=False)
    fs = FlowSolver3D(cfg)
    rng = np.random.default_rng(31)
    ops = fs.ops
    u = torch.as_tensor(0.3 * rng.standard_normal((ops.nz, ops.ny, ops.nx + 1)))
    v = torch.as_tensor(0.3 * rng.standard_normal((ops.nz, ops.ny + 1, ops.nx)))
    w = torch.as_tensor(0.3 * rng.standard_normal((ops.nz + 1, ops.ny, ops.nx)))
    u, v, w = core3.apply_boundary_values(u, v, w, ops)
    u, v, w, _ = core3.project3d(u, v, w, ops)
    st = fs.state
    with torch.no_grad():
        st.u.copy_(u)
        st.v.copy_(v)
        st.w.copy_(w)
    arrs = {"start:u": np_(u), "start:v": np_(v), "start:w": np_(w)}
    fs.advance(6)
    arrs.update(flatten(fs.state_dict()))
    fs.advance(4)
    st = fs.state
    arrs.update({"after:u": np_(st.u), "after:v": np_(st.v), "after:w": np_(st.w), "after:p": np_(st.p),
                 "after:t": np_(st.t)})
    np.savez_compressed(OUT / "ckpt_ns3d.npz", **arrs)


if __name__ == "__main__":
    torch.set_num_threads(8)
    fixture_2d()
    fixture_3d()
    print("written", flush=True)
