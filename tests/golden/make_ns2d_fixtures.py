"""Write the 2D golden fixtures from the reference solver (run in the build container only).

The reference is imported from ``/root/reference/src`` (read-only) and run on the CPU. The
fixtures pin what no closed form pins: the composition of a whole step (stage order,
inlet/outlet correction before penalisation, buoyancy after it, step-averaged pressure),
the mask builder, the emitter footprints and the thermal heat accounting.

Each file holds inputs and expected outputs only (numpy arrays, no code):

* ``ns2d_trace_64x32.npz`` - one traced float64 step of a 64x32 tunnel with a disc, with
  every stage field, plus 20 more steps.
* ``ns2d_c1_f32.npz`` / ``ns2d_c1_f64.npz`` - the 256x128 disc case (Re 100, upwind3, no LES)
  from a seeded noisy start: u, v, p at steps 1, 10, 100 and 500, dt and [Cd, Cl, Cm] per step.
* ``ns2d_c2mini_f64.npz`` / ``ns2d_c2mini_f32.npz`` - a 128x64 miniature of the heated-airfoil
  case (Smagorinsky, NACA 0012 at 8 deg held at theta = 1, Pr 0.71, Ri 0.1, a line smoke
  emitter), 20 steps, with theta, dye and the per-step heat outputs.
* ``ns2d_masks.npz`` - mask tensors of the 256x128 disc and of the 1024x512 airfoil.

Usage: ``python tests/golden/make_ns2d_fixtures.py`` (about a minute on 8 cores).
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

from aerolume.solvers.ns2d import core, physics  # noqa: E402
from aerolume.solvers.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig  # noqa: E402
from aerolume.solvers.ns2d.geometry import build_geometry  # noqa: E402
from aerolume.solvers.ns2d.solver import FlowSolver  # noqa: E402


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def naca4(code="0012", n=400, chord=1.0):
    t = int(code[2:]) / 100
    beta = np.linspace(0.0, np.pi, n // 2 + 1)
    x = 0.5 * (1 - np.cos(beta))
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 - 0.1036 * x**4)
    upper = np.c_[x, yt][::-1]
    lower = np.c_[x, -yt][1:-1]
    return np.vstack([upper, lower]) * chord


def rotate(pts, deg, about=(0.0, 0.0)):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    p = np.asarray(pts, float) - about
    return np.c_[c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1]] + about


def noisy_start(ops, seed):
    """Inlet flow + 1 % Gaussian noise on every face, then boundary values and projection."""
    rng = np.random.default_rng(seed)
    nu_ = rng.standard_normal((ops.ny, ops.nx + 1))
    nv_ = rng.standard_normal((ops.ny + 1, ops.nx))
    u = torch.as_tensor(1.0 + 0.01 * nu_, dtype=ops.dtype)
    v = torch.as_tensor(0.01 * nv_, dtype=ops.dtype)
    u, v = core._apply_boundary_values(u, v, ops)
    u, v, _ = core.project(u, v, ops)
    return u, v


def geom_arrays(geom, prefix="g_"):
    out = {prefix + "window": np.asarray(geom.window, dtype=np.int64), prefix + "nbody": np.int64(geom.nbody)}
    for k in ("chi_u", "chi_v", "chi_c", "a_u", "a_v", "k_u", "k_v"):
        out[prefix + k] = getattr(geom, k).detach().cpu().numpy()
    tg = geom.thermal
    if tg is not None:
        for k in ("chi_t", "a_t", "w_t", "ins_u", "ins_v", "q_skin", "w_f", "theta_s", "flux_g", "perimeter",
                  "vol_u", "vol_v"):
            out[prefix + "th_" + k] = getattr(tg, k).detach().cpu().numpy()
    return out


def np_(t):
    return t.detach().cpu().numpy()


def trace_fixture():
    cfg = SolverConfig(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=100.0,
                       bodies=[ImmersedBody(circle(64, 0.75, 2.0, 2.0))], advection="upwind3",
                       turbulence="none", cfl=0.5, dtype="float64", device="cpu", dye=False, perturb=False)
    ops = core.make_operators(cfg)
    geom = build_geometry(cfg.bodies, ops)
    u, v = noisy_start(ops, 7)
    st = core.init_state(cfg, ops)._replace(u=u, v=v)
    tr = {}
    new, coeffs = core.step_full(st, geom, ops, None, trace=tr)
    arrs = {"u0": np_(u), "v0": np_(v), "u1": np_(new.u), "v1": np_(new.v), "p1": np_(new.p),
            "dt1": np_(new.dt), "coeffs1": np_(coeffs)}
    for k, t in tr.items():
        arrs["trace:" + k] = np_(t)
    cur = new
    cs = []
    for _ in range(20):
        cur, c = core.step_full(cur, geom, ops, None)
        cs.append(np_(c))
    arrs.update({"u21": np_(cur.u), "v21": np_(cur.v), "p21": np_(cur.p), "t21": np_(cur.t),
                 "coeffs_hist": np.stack(cs)})
    arrs.update(geom_arrays(geom))
    np.savez_compressed(OUT / "ns2d_trace_64x32.npz", **arrs)


def c1_fixture(dtype):
    cfg = SolverConfig(nx=256, ny=128, h=1.0 / 32.0, x0=0.0, y0=0.0, re=100.0,
                       bodies=[ImmersedBody(circle(256, 0.5, 2.0, 2.0))], advection="upwind3",
                       turbulence="none", cfl=0.5, dtype=dtype, device="cpu", dye=False, perturb=False)
    ops = core.make_operators(cfg)
    geom = build_geometry(cfg.bodies, ops)
    u, v = noisy_start(ops, 101)
    st = core.init_state(cfg, ops)._replace(u=u, v=v)
    arrs = {"u0": np_(u), "v0": np_(v)}
    dts, cs = [], []
    cur = st
    keep = (1, 10, 100, 500) if dtype == "float32" else (10, 500)
    for n in range(1, 501):
        cur, c = core.step_full(cur, geom, ops, None)
        dts.append(float(cur.dt))
        cs.append(np_(c)[0])
        if n in keep:
            arrs[f"u{n}"], arrs[f"v{n}"], arrs[f"p{n}"] = np_(cur.u), np_(cur.v), np_(cur.p)
    arrs["dt_hist"] = np.asarray(dts)
    arrs["coeffs_hist"] = np.stack(cs)
    arrs["t500"] = np_(cur.t)
    if dtype == "float32":
        arrs.update(geom_arrays(geom))
    tag = "f32" if dtype == "float32" else "f64"
    np.savez_compressed(OUT / f"ns2d_c1_{tag}.npz", **arrs)


def c2_config(nx, ny, dtype):
    h = 8.0 / nx
    foil = rotate(naca4("0012", 400), -8.0, about=(0.25, 0.0))
    return SolverConfig(nx=nx, ny=ny, h=h, x0=-2.0, y0=-2.0, re=1e5,
                        bodies=[ImmersedBody(foil, thermal="temperature", thermal_value=1.0)],
                        advection="upwind3", turbulence="smagorinsky", cfl=0.5, dtype=dtype, device="cpu",
                        dye=True, dye_inlet=False, perturb=False,
                        thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                        emitters=[EmitterSpec(kind="line", position=(-1.0, -0.2), end=(-1.0, 0.2), channel=0,
                                              rate=1.0)])


def c2mini_fixture(dtype):
    cfg = c2_config(128, 64, dtype)
    solver = FlowSolver(cfg)
    ops, geom = solver.ops, solver.geometry
    u, v = noisy_start(ops, 202)
    st = solver.state
    st0 = st._replace(u=u, v=v)
    ph = ops.phys
    arrs = {"u0": np_(u), "v0": np_(v), "theta0": np_(st.theta), "dye0": np_(st.dye)}
    cur = st0
    rows = {k: [] for k in ("heat", "wall_heat", "heat_storage", "flow_heat", "coeffs", "dt", "t")}
    for n in range(1, 21):
        out = {}
        cur, c = core.step_full(cur, geom, ops, None, out=out)
        rows["coeffs"].append(np_(c)[0])
        rows["dt"].append(float(cur.dt))
        rows["t"].append(float(cur.t))
        for k in ("heat", "wall_heat", "heat_storage", "flow_heat"):
            rows[k].append(np_(out[k]))
        if n in (1, 20):
            for k in ("u", "v", "p", "theta", "dye"):
                arrs[f"{k}{n}"] = np_(getattr(cur, k))
    for k, r in rows.items():
        arrs["hist_" + k] = np.asarray(r)
    arrs.update(geom_arrays(geom))
    g = ph.groups[0]
    arrs.update({"dye_weight": np_(g.weight), "dye_target": np_(g.target), "dye_timing": np_(g.timing),
                 "kappa": np_(ph.kappa), "ri": np_(ph.ri), "eta_t": np_(ph.eta_t), "eta": np_(ops.eta),
                 "diff_coef": np.float64(ph.diff_coef)})
    tag = "f32" if dtype == "float32" else "f64"
    np.savez_compressed(OUT / f"ns2d_c2mini_{tag}.npz", **arrs)


def masks_fixture():
    arrs = {}
    cfg1 = SolverConfig(nx=256, ny=128, h=1.0 / 32.0, x0=0.0, y0=0.0, re=100.0,
                        bodies=[ImmersedBody(circle(256, 0.5, 2.0, 2.0))], dtype="float64", device="cpu",
                        dye=False, perturb=False)
    ops1 = core.make_operators(cfg1)
    arrs.update(geom_arrays(build_geometry(cfg1.bodies, ops1), prefix="c1_"))
    cfg2 = c2_config(1024, 512, "float64")
    s2 = FlowSolver(cfg2)
    arrs.update(geom_arrays(s2.geometry, prefix="c2_"))
    w = s2.ops.phys.groups[0].weight
    nz = torch.nonzero(w > 0)
    j0, j1 = int(nz[:, 0].min()), int(nz[:, 0].max()) + 1
    i0, i1 = int(nz[:, 1].min()), int(nz[:, 1].max()) + 1
    arrs["c2_dye_box"] = np.asarray([j0, j1, i0, i1], dtype=np.int64)
    arrs["c2_dye_weight"] = np_(w[j0:j1, i0:i1])
    np.savez_compressed(OUT / "ns2d_masks.npz", **arrs)


if __name__ == "__main__":
    torch.set_num_threads(8)
    torch.set_flush_denormal(True)
    trace_fixture()
    print("trace done", flush=True)
    masks_fixture()
    print("masks done", flush=True)
    c2mini_fixture("float64")
    c2mini_fixture("float32")
    print("c2mini done", flush=True)
    c1_fixture("float32")
    c1_fixture("float64")
    print("c1 done", flush=True)
