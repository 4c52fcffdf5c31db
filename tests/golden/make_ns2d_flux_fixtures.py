"""Write the golden fixture of 2D heat-flux bodies from the reference solver (build container
only: it imports the reference from ``/root/reference/src``, read-only, and runs it on the CPU).

``ns2d_flux_bodies.npz`` (float64): the two-disc tunnel of ``tests/cases.py`` ``flux_cfg`` -
a disc releasing heat at G = 2 through its skin upstream of a disc held at theta = 1 - from a
seeded noisy start, 20 steps of ``core.step_full``. Holds the reference's thermal window
arrays (``q_skin``, ``w_f``, ``perimeter``, ``ins_u``, ``ins_v``, ``chi_t``), the start, u, v,
p and theta after steps 1 and 20, and per step dt, t, the coefficients and the heat outputs
(``heat`` of the held disc, ``theta_wall`` of the heat-flux disc, ``wall_heat``,
``heat_storage``, ``flow_heat``).

Usage: ``python tests/golden/make_ns2d_flux_fixtures.py`` (seconds).
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
from aerolume.solvers.ns2d.config import ImmersedBody, SolverConfig, ThermalConfig  # noqa: E402
from aerolume.solvers.ns2d.solver import FlowSolver  # noqa: E402


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def np_(t):
    return t.detach().cpu().numpy()


def main():
    torch.set_num_threads(8)
    cfg = SolverConfig(nx=128, ny=64, h=1.0 / 16.0, x0=-2.0, y0=-2.0, re=2000.0,
                       bodies=[ImmersedBody(circle(96, 0.6, 0.0, 0.4), name="flux", thermal="heat_flux",
                                            thermal_value=2.0),
                               ImmersedBody(circle(96, 0.5, 1.5, -0.5), name="held", thermal="temperature",
                                            thermal_value=1.0)],
                       advection="upwind3", turbulence="smagorinsky", cfl=0.5, dtype="float64", device="cpu",
                       dye=False, perturb=False, thermal=ThermalConfig(prandtl=0.71, richardson=0.1))
    solver = FlowSolver(cfg)
    ops, geom = solver.ops, solver.geometry
    tg = geom.thermal
    rng = np.random.default_rng(505)
    u = torch.as_tensor(1.0 + 0.01 * rng.standard_normal((ops.ny, ops.nx + 1))) * ops.u_inlet.expand(ops.ny, ops.nx + 1)
    v = torch.as_tensor(0.01 * rng.standard_normal((ops.ny + 1, ops.nx)))
    u, v = core._apply_boundary_values(u, v, ops)
    u, v, _ = core.project(u, v, ops)
    st = solver.state._replace(u=u, v=v)
    arrs = {"u0": np_(u), "v0": np_(v), "theta0": np_(st.theta), "window": np.asarray(geom.window, np.int64),
            "flux_ids": np.asarray(tg.flux_ids, np.int64), "temp_ids": np.asarray(tg.temp_ids, np.int64)}
    for k in ("q_skin", "w_f", "perimeter", "ins_u", "ins_v", "chi_t", "flux_g"):
        arrs["g_" + k] = np_(getattr(tg, k))
    rows = {k: [] for k in ("heat", "theta_wall", "wall_heat", "heat_storage", "flow_heat", "coeffs", "dt", "t")}
    cur = st
    with torch.no_grad():
        for n in range(1, 21):
            out = {}
            cur, c = core.step_full(cur, geom, ops, None, out=out)
            rows["coeffs"].append(np_(c))
            rows["dt"].append(float(cur.dt))
            rows["t"].append(float(cur.t))
            for k in ("heat", "theta_wall", "wall_heat", "heat_storage", "flow_heat"):
                rows[k].append(np_(out[k]))
            if n in (1, 20):
                for k in ("u", "v", "p", "theta"):
                    arrs[f"{k}{n}"] = np_(getattr(cur, k))
    for k, r in rows.items():
        arrs["hist_" + k] = np.asarray(r)
    np.savez_compressed(OUT / "ns2d_flux_bodies.npz", **arrs)
    print("written", flush=True)


if __name__ == "__main__":
    main()
