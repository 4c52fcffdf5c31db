"""Write the golden fixtures of the reference's default 2D path and of the 2D variants the
other fixtures leave out (run in the build container only; it imports the reference from
``/root/reference/src``, read-only, and runs it on the CPU).

* ``ns2d_default_f64.npz`` / ``ns2d_default_f32.npz`` - ``FlowSolver(cfg)`` with every default
  the reference's ``SolverConfig`` sets (``dye=True``, ``dye_inlet=True``, ``perturb=True``,
  upwind3, no LES): a 128x64 tunnel (h = 1/16, Re 200) with a disc of diameter 0.5 at (2, 2).
  The initial state (the Lamb-Oseen start, ns2d/core.py:216-271) and the state after 1, 10
  and 20 steps (u, v, p, the inlet smoke comb's dye, t, dt) plus the coefficient history.
* ``ns2d_variants.npz`` - five float64 steps of ``core.step_full`` from a seeded noisy start
  for the 2D features no other fixture steps: central advection with Vreman and no-slip
  walls behind a parabolic inlet; WALE over a moving bottom wall; a doubly periodic box with
  Smagorinsky; a closed lid-driven box with slip side walls; the Werner-Wengle wall model
  (``wall_model="werner_wengle"``) in a Re 1e5 tunnel over a moving floor under a no-slip
  roof, and in a closed Re 1e6 room with a moving floor and no-slip walls all round.

Each file holds inputs and expected outputs only.

Usage: ``python tests/golden/make_ns2d_default_fixtures.py [variants]`` (about half a minute;
``variants``: only ``ns2d_variants.npz``).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent
sys.path.insert(0, str(REF))

import torch  # noqa: E402

from aerolume.solvers.ns2d import core  # noqa: E402
from aerolume.solvers.ns2d.config import ImmersedBody, SolverConfig  # noqa: E402
from aerolume.solvers.ns2d.geometry import build_geometry  # noqa: E402
from aerolume.solvers.ns2d.solver import FlowSolver  # noqa: E402


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def np_(t):
    return t.detach().cpu().numpy()


def default_fixture(dtype):
    cfg = SolverConfig(nx=128, ny=64, h=1.0 / 16.0, x0=0.0, y0=0.0, re=200.0,
                       bodies=[ImmersedBody(circle(128, 0.5, 2.0, 2.0))], dtype=dtype, device="cpu")
    assert cfg.dye and cfg.dye_inlet and cfg.perturb
    fs = FlowSolver(cfg)
    arrs = {}

    def snap(tag):
        st = fs.state
        arrs.update({f"u_{tag}": np_(st.u), f"v_{tag}": np_(st.v), f"p_{tag}": np_(st.p),
                     f"dye_{tag}": np_(st.dye), f"t_{tag}": np_(st.t), f"dt_{tag}": np_(st.dt)})
    snap(0)
    done = 0
    for n in (1, 10, 20):
        fs.advance(n - done)
        done = n
        snap(n)
    hist = fs.history()
    for k, v in hist.items():
        arrs["hist:" + k] = np.asarray(v)
    arrs["coefficients"] = np.asarray(fs.coefficients())
    np.savez_compressed(OUT / f"ns2d_default_{'f64' if dtype == 'float64' else 'f32'}.npz", **arrs)


def noisy_start(ops, seed, closed=False, periodic=False):
    rng = np.random.default_rng(seed)
    base = 0.0 if (closed or periodic) else 1.0
    u = torch.as_tensor(base + 0.01 * rng.standard_normal((ops.ny, ops.nx + 1)), dtype=ops.dtype)
    v = torch.as_tensor(0.01 * rng.standard_normal((ops.ny + 1, ops.nx)), dtype=ops.dtype)
    if not (closed or periodic):
        u = u * ops.u_inlet.expand_as(u)
    u, v = core._apply_boundary_values(u, v, ops)
    u, v, _ = core.project(u, v, ops)
    return u, v


VARIANTS = {
    "central_vreman_noslip": dict(advection="central", turbulence="vreman", bc_top="noslip", bc_bottom="noslip",
                                  inlet="parabolic", bodies=True),
    "wale_moving": dict(turbulence="wale", bc_bottom="moving", ground_speed=1.0, bodies=True),
    "periodic_smag": dict(periodic=True, turbulence="smagorinsky", bodies=False),
    "closed_lid": dict(domain="closed", bc_left="slip", bc_right="slip", bc_bottom="moving", ground_speed=1.0,
                       bodies=True),
    "ww_tunnel_moving": dict(re=1e5, turbulence="smagorinsky", bc_bottom="moving", ground_speed=1.0,
                             bc_top="noslip", wall_model="werner_wengle", bodies=True),
    "ww_closed_room": dict(re=1e6, domain="closed", bc_left="noslip", bc_right="noslip", bc_bottom="moving",
                           ground_speed=1.0, bc_top="noslip", wall_model="werner_wengle", bodies=True),
}


def variants_fixture():
    arrs = {}
    for name, kw in VARIANTS.items():
        kw = dict(kw)
        bodies = [ImmersedBody(circle(64, 0.75, 2.0, 2.0))] if kw.pop("bodies") else []
        cfg = SolverConfig(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=kw.pop("re", 400.0), bodies=bodies,
                           dtype="float64", device="cpu", dye=False, perturb=False, **kw)
        ops = core.make_operators(cfg)
        geom = build_geometry(cfg.bodies, ops)
        u, v = noisy_start(ops, 17, closed=kw.get("domain") == "closed", periodic=kw.get("periodic", False))
        st = core.init_state(cfg, ops)._replace(u=u, v=v)
        arrs[f"{name}:u0"], arrs[f"{name}:v0"] = np_(u), np_(v)
        cs = []
        for _ in range(5):
            st, c = core.step_full(st, geom, ops, None)
            cs.append(np_(c))
        arrs.update({f"{name}:u5": np_(st.u), f"{name}:v5": np_(st.v), f"{name}:p5": np_(st.p),
                     f"{name}:t5": np_(st.t), f"{name}:coeffs": np.stack(cs) if cs[0].size else np.zeros((5, 0, 3))})
    np.savez_compressed(OUT / "ns2d_variants.npz", **arrs)


if __name__ == "__main__":
    torch.set_num_threads(8)
    if sys.argv[1:] != ["variants"]:
        default_fixture("float64")
        default_fixture("float32")
    variants_fixture()
    print("written", flush=True)
