"""Write the 2D PID-step and diagnostics golden fixture from the reference solver (build
container only).

The reference is imported from ``/root/reference/src`` (read-only) and run on the CPU in
float64. ``ns2d_pid_diag.npz`` holds inputs and expected outputs only (numpy arrays):

* the start of a 64x32 tunnel with a heated disc (Smagorinsky, Pr 0.71, Ri 0.1,
  ``adaptive_dt="pid"``): u, v after the reference's boundary values and projection, theta;
* ``core.step_full`` from that start for 30 steps: dt and the controller memory ``ctrl``
  after every step, u, v, p, theta after the last;
* the reference ``FlowSolver`` with ``diagnostics=True`` on the same case and start, advanced
  7 then 13 steps: its diagnostics rows (one per 10-step block, columns ``DIAG_COLUMNS``)
  and the state after 20 steps;
* the same solver's host reports after those 20 steps, with running means started after step
  7: ``surface(48)`` and ``surface(48, mean=True)`` (``surf_`` / ``msurf_`` + key),
  ``heat_transfer()`` / ``heat_balance()`` and their ``mean=True`` forms (``ht_`` / ``mht_``,
  ``hb_`` / ``mhb_``), ``mean_fields()`` (``mean_`` + key) and ``averaging_time``.

``ns2d_particles.npz``: the same case and start with two particle emitters (a continuous
5-nozzle rake, rate 400, and a pulsed 6-nozzle line, period 0.15) on a 512-particle pool, the
reference ``FlowSolver`` advanced 25 steps: the pool's buffers (pos, birth, alive, head, acc,
last), the snapshot and the number of polylines.

Usage: ``python tests/golden/make_ns2d_diag_fixtures.py`` (a few seconds).
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
from aerolume.solvers.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig  # noqa: E402
from aerolume.solvers.ns2d.diagnostics import DIAG_COLUMNS  # noqa: E402
from aerolume.solvers.ns2d.geometry import build_geometry  # noqa: E402
from aerolume.solvers.ns2d.solver import FlowSolver  # noqa: E402

SEED = 808


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def config(**over):
    base = dict(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=400.0,
                bodies=[ImmersedBody(circle(64, 0.75, 2.0, 2.0), thermal="temperature", thermal_value=1.0)],
                advection="upwind3", turbulence="smagorinsky", cfl=0.5, dtype="float64", device="cpu",
                dye=False, perturb=False, thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                adaptive_dt="pid", use_cuda_graphs=False)
    base.update(over)
    return SolverConfig(**base)


def start(ops):
    rng = np.random.default_rng(SEED)
    u = torch.as_tensor(1.0 + 0.01 * rng.standard_normal((ops.ny, ops.nx + 1)), dtype=ops.dtype)
    v = torch.as_tensor(0.01 * rng.standard_normal((ops.ny + 1, ops.nx)), dtype=ops.dtype)
    u, v = core._apply_boundary_values(u, v, ops)
    u, v, _ = core.project(u, v, ops)
    theta = torch.as_tensor(0.2 * rng.random((ops.ny, ops.nx)), dtype=ops.dtype)
    return u, v, theta


def np_(t):
    return t.detach().cpu().numpy()


def main():
    torch.set_num_threads(8)
    cfg = config()
    ops = core.make_operators(cfg)
    geom = build_geometry(cfg.bodies, ops)
    u, v, theta = start(ops)
    st = core.init_state(cfg, ops)._replace(u=u, v=v, theta=theta)
    arrs = {"u0": np_(u), "v0": np_(v), "theta0": np_(theta), "ctrl0": np_(st.ctrl)}
    dts, ctrls = [], []
    cur = st
    with torch.no_grad():
        for _ in range(30):
            cur, _ = core.step_full(cur, geom, ops, None)
            dts.append(float(cur.dt))
            ctrls.append(np_(cur.ctrl).copy())
    arrs.update(dt_hist=np.asarray(dts), ctrl_hist=np.stack(ctrls), u30=np_(cur.u), v30=np_(cur.v),
                p30=np_(cur.p), theta30=np_(cur.theta))
    s = FlowSolver(config(diagnostics=True))
    s._state = s._state._replace(u=u.clone(), v=v.clone(), theta=theta.clone())
    s.advance(7)
    s.start_averaging()
    s.advance(13)
    d = s.diagnostics()
    for pre, mean in (("surf_", False), ("msurf_", True)):
        rec = s.surface(48, mean=mean)[0]
        for k, val in rec.items():
            if k != "name":
                arrs[pre + k] = np.asarray(val, dtype=np.float64)
    for pre, mean in (("ht_", False), ("mht_", True)):
        ht = s.heat_transfer(mean=mean)
        b = ht["bodies"][0]
        for k in ("perimeter", "heat_rate", "theta_wall", "nusselt"):
            arrs[pre + k] = np.float64(b[k])
        arrs[pre + "walls"] = np.array([ht["walls"][w] for w in ("left", "right", "bottom", "top")])
        arrs[pre + "nusselt_mean"] = np.float64(ht["nusselt_mean"])
        arrs[pre + "pe"] = np.float64(ht["pe"])
    for pre, mean in (("hb_", False), ("mhb_", True)):
        hb = s.heat_balance(mean=mean)
        arrs[pre + "keys"] = np.array(list(hb))
        arrs[pre + "values"] = np.array([hb[k] for k in hb], dtype=np.float64)
    for k, val in s.mean_fields().items():
        arrs["mean_" + k] = np_(val)
    arrs["averaging_time"] = np.float64(s.averaging_time)
    for k in DIAG_COLUMNS:
        arrs["diag_" + k] = np.asarray(d[k])
    arrs["diag_columns"] = np.array(DIAG_COLUMNS)
    st20 = s.state
    arrs.update(u20=np_(st20.u), v20=np_(st20.v), p20=np_(st20.p), theta20=np_(st20.theta),
                dt20=np.float64(float(st20.dt)), ctrl20=np_(st20.ctrl))
    np.savez_compressed(OUT / "ns2d_pid_diag.npz", **arrs)
    print("wrote ns2d_pid_diag.npz", dts[:3], d["step"], flush=True)


PARTICLE_EMITTERS = dict(
    rake=dict(position=(0.6, 1.2), kind="rake", end=(0.6, 2.8), count=5, species="particles", rate=400.0),
    line=dict(position=(1.2, 1.0), kind="line", end=(1.2, 3.0), count=6, species="particles", rate=1.0,
              pulse_period=0.15, pulse_duty=0.5, start=0.05))


def particles_case():
    ems = [EmitterSpec(**PARTICLE_EMITTERS["rake"]), EmitterSpec(**PARTICLE_EMITTERS["line"])]
    cfg = config(emitters=ems, particle_capacity=512)
    ops = core.make_operators(cfg)
    u, v, theta = start(ops)
    s = FlowSolver(cfg)
    s._state = s._state._replace(u=u.clone(), v=v.clone(), theta=theta.clone())
    s.advance(25)
    pt = s._particles
    arrs = {k: np_(getattr(pt, k)).copy() for k in ("pos", "birth", "alive", "head", "acc", "last")}
    snap = s.particles()
    for k, val in snap.items():
        arrs["snap_" + k] = np.asarray(val)
    arrs["n_lines"] = np.int64(len(s.particle_lines()))
    arrs["t"] = np.float64(s.t)
    np.savez_compressed(OUT / "ns2d_particles.npz", **arrs)
    print("wrote ns2d_particles.npz", int(arrs["alive"].sum()), "alive", flush=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["particles"]:
        torch.set_num_threads(8)
        particles_case()
        sys.exit(0)
    main()
