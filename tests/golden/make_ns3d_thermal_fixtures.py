"""Write the 3D heat-transfer golden fixtures from the reference solver (build container only).

The reference is imported from ``/root/reference/src`` (read-only) and run on the CPU in
float64: ``core.step3d_phys`` with a ``physics.Physics3D`` that carries heat transfer only.
Its mask builders run with their own signed distance (the optional shared voxeliser is
switched off, as the 3D mask fixtures do), on triangulated spheres and boxes.

Each ``ns3d_thermal_<case>.npz`` holds inputs and expected outputs only (numpy arrays):
the meshes (vertices, faces per body), the reference's thermal masks (chi_t, a_t, w_t,
adiabatic face fractions, vol_w, the heat-flux skin and its cell list, flux rates), the side
kinds, the seeded start (u, v, w after the reference's projection, theta), u, v, w, p and
theta after the first and the last step, dt, t, the six coefficients and the bodies' heat
rates of every step, the wall heat rates and wall temperatures at the end.

``ns3d_diag_box.npz``: the reference's ``FlowSolver3D`` with ``diagnostics=True`` on the "box"
case from the same start (``load_state_dict``), advanced 7 then 13 steps with running means
started after step 7: its diagnostics rows (one per 10-step block), the state after 20 steps,
``mean_fields()`` (``mean_`` + key), ``averaging_time``, slices (``slice_<axis>_<name>`` and
``mslice_`` of the means), ``surface()`` / ``surface(mean=True)`` of body 0 (``surf_`` /
``msurf_`` + key) and ``heat_balance()`` (``hb_keys``, ``hb_values``).

``ns3d_pid_box.npz``: ``core.step3d_phys`` on the "box" case from the same start with
``adaptive_dt="pid"`` (``pid_target`` 0.8) for 30 steps: dt and the controller memory
(dt_prev, e1, e2) after every step, u, v, w, p and theta after the last.

Usage: ``python tests/golden/make_ns3d_thermal_fixtures.py [pid]`` (about a minute on 8 cores;
``pid`` / ``diag``: only that fixture).
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent
ROOT = OUT.parent.parent
sys.path.insert(0, str(REF))
sys.path.insert(0, str(ROOT))

import torch  # noqa: E402

import aerolume.solvers.ns3d.geometry as rgeo  # noqa: E402
from aerolume.solvers.ns3d import core as rcore  # noqa: E402
from aerolume.solvers.ns3d.config import HeatSource3D as RHeat  # noqa: E402
from aerolume.solvers.ns3d.config import ImmersedMesh as RMesh  # noqa: E402
from aerolume.solvers.ns3d.config import SolverConfig3D as RConfig  # noqa: E402
from aerolume.solvers.ns3d.config import ThermalConfig3D as RThermal  # noqa: E402
from aerolume.solvers.ns3d.physics import Physics3D  # noqa: E402

from tests.cases import THERMAL3D as CASES  # noqa: E402
from tests.cases import THERMAL_SEED as SEED  # noqa: E402

rgeo._SHARED = False          # the reference's own torch signed distance, no shared module


def _cases_module():
    path = Path("/root/reference/validation/ns3d_cases.py")
    spec = importlib.util.spec_from_file_location("ref_ns3d_cases", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ref_ns3d_cases"] = mod
    spec.loader.exec_module(mod)
    return mod


CM = _cases_module()
GEOM_KEYS = ("chi_u", "chi_v", "chi_w", "chi_c", "a_u", "a_v", "a_w", "k_u", "k_v", "k_w")



def meshes(spec):
    out = []
    for shape, size, centre, mode, value in spec["bodies"]:
        if shape == "sphere":
            v, f = CM.icosphere(size, 2, centre)
        else:
            v, f = CM.box_mesh(*size, center=centre, n=2)
        out.append((np.asarray(v, np.float64), np.asarray(f, np.int64), mode, value))
    return out


def run(name, spec):
    bodies = [RMesh(v, f, name=f"b{i}", thermal=m, thermal_value=val)
              for i, (v, f, m, val) in enumerate(meshes(spec))]
    cfg = RConfig(**spec["grid"], bodies=bodies, device="cpu", dtype="float64", dye=False, perturb=False,
                  ref_area=0.3, length_ref=0.8, span_ref=0.6, moment_ref=(0.2, 0.0, 0.5),
                  thermal=RThermal(**spec["thermal"]), heat_sources=[RHeat(**h) for h in spec["heaters"]])
    ops = rcore.make_operators3d(cfg)
    g = rgeo.build_geometry3d(bodies, ops, use_shared=False)
    phys = Physics3D(cfg, ops, g, bodies)
    th = phys.thermal
    rng = np.random.default_rng(SEED)
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    base = 0.0 if (all(ops.periodic_axes) or ops.closed) else 1.0
    u = torch.as_tensor(base + 0.05 * rng.standard_normal((nz, ny, nx + 1)))
    v = torch.as_tensor(0.05 * rng.standard_normal((nz, ny + 1, nx)))
    w = torch.as_tensor(0.05 * rng.standard_normal((nz + 1, ny, nx)))
    u, v, w = rcore.apply_boundary_values(u, v, w, ops)
    u, v, w, _ = rcore.project3d(u, v, w, ops)
    theta = torch.as_tensor(0.3 * rng.random((nz, ny, nx)))
    out = {"u0": u.numpy(), "v0": v.numpy(), "w0": w.numpy(), "theta0": theta.numpy().copy(),
           "window": np.asarray(g.window, np.int64), "side_kinds": np.array(th.side_kinds),
           "side_vals": th.side_vals.numpy(), "q_max": np.float64(th.q_max), "eta_t": th.eta.numpy(),
           "source_total": np.float64(th.sources.total), "flux_rate": th.geom.flux_rate.numpy(),
           "area": th.geom.area}
    for i, (vv, ff, m, val) in enumerate(meshes(spec)):
        out[f"mesh{i}_v"], out[f"mesh{i}_f"] = vv, ff
    for k in GEOM_KEYS:
        out["g_" + k] = getattr(g, k).numpy()
    tg = th.geom
    for k in ("chi_t", "a_t", "w_t", "vol_w", "skin", "wall_idx", "wall_w", "wall_seg"):
        val = getattr(tg, k)
        if val is not None:
            out["t_" + k] = val.numpy()
    if tg.adia is not None:
        for c, a in zip("uvw", tg.adia):
            out["t_adia_" + c] = a.numpy()
    for i, (win, q) in enumerate(th.sources.items):
        out[f"src{i}_window"], out[f"src{i}_q"] = np.asarray(win, np.int64), q.numpy()
    z = torch.zeros((), dtype=torch.float64)
    st = rcore.State3D(u, v, w, torch.zeros(nz, ny, nx, dtype=torch.float64), z.clone(), z.clone())
    dts, ts, coeffs, heat = [], [], [], []
    steps = spec["steps"]
    with torch.no_grad():
        for n in range(1, steps + 1):
            r = rcore.step3d_phys(st, g, ops, phys=phys, theta=theta)
            st, theta = r.state, r.theta
            dts.append(float(st.dt))
            ts.append(float(st.t))
            coeffs.append(r.coeffs.numpy().copy())
            heat.append(th.heat_rate.numpy().copy())
            if n in (1, steps):
                for f in ("u", "v", "w", "p"):
                    out[f"{f}_{n}"] = getattr(st, f).numpy().copy()
                out[f"theta_{n}"] = theta.numpy().copy()
        walls = th.wall_heat_rates(theta, st, ops)
    out["steps"] = np.int64(steps)
    out["dt"], out["t"] = np.array(dts), np.array(ts)
    out["coeffs"], out["heat_rate"] = np.stack(coeffs), np.stack(heat)
    out["wall_heat"] = np.array([walls[s] for s in ("left", "right", "front", "back", "bottom", "top")])
    out["wall_theta"] = tg.wall_theta(theta).numpy()
    return out


def diag_case():
    from aerolume.solvers.ns3d.solver import DIAG_KEYS
    from aerolume.solvers.ns3d.solver import FlowSolver3D as RSolver
    spec = CASES["box"]
    fx = dict(np.load(OUT / "ns3d_thermal_box.npz"))
    bodies = [RMesh(fx[f"mesh{i}_v"], fx[f"mesh{i}_f"], name=f"b{i}", thermal=m, thermal_value=val)
              for i, (_, _, _, m, val) in enumerate(spec["bodies"])]
    cfg = RConfig(**spec["grid"], bodies=bodies, device="cpu", dtype="float64", dye=False, perturb=False,
                  ref_area=0.3, length_ref=0.8, span_ref=0.6, moment_ref=(0.2, 0.0, 0.5),
                  thermal=RThermal(**spec["thermal"]), heat_sources=[RHeat(**h) for h in spec["heaters"]],
                  diagnostics=True, use_cuda_graphs=False, fused_kernels=False)
    s = RSolver(cfg)
    n = s.ops
    z = torch.zeros((), dtype=torch.float64)
    s.load_state_dict({"u": torch.as_tensor(fx["u0"]), "v": torch.as_tensor(fx["v0"]), "w": torch.as_tensor(fx["w0"]),
                       "p": torch.zeros(n.nz, n.ny, n.nx, dtype=torch.float64), "t": z.clone(), "dt": z.clone(),
                       "theta": torch.as_tensor(fx["theta0"]), "steps": 0})
    with torch.no_grad():
        s.advance(7)
        s.start_averaging()
        s.advance(13)
    d = s.diagnostics()
    out = {k: np.asarray(d[k]) for k in DIAG_KEYS}
    with torch.no_grad():
        for k, v in s.mean_fields().items():
            out["mean_" + k] = v.numpy().copy()
        out["averaging_time"] = np.float64(s.averaging_time)
        for axis, pos in (("x", 0.37), ("y", 0.5), ("z", 0.21)):
            for name in ("u", "p", "temperature", "q"):
                out[f"slice_{axis}_{name}"] = s.slice(axis, pos, name).numpy().copy()
            for name in ("w", "u_rms", "temperature_rms"):
                out[f"mslice_{axis}_{name}"] = s.slice(axis, pos, name, mean=True).numpy().copy()
        for pre, mean in (("surf_", False), ("msurf_", True)):
            rec = s.surface(mean=mean)[0]
            for k, v in rec.items():
                if k != "name":
                    out[pre + k] = np.asarray(v)
        hb = s.heat_balance()
        out["hb_keys"] = np.array(list(hb))
        out["hb_values"] = np.array([hb[k] for k in hb], dtype=np.float64)
    out["step"] = np.asarray(d["step"])
    st = s.state
    for f in ("u", "v", "w", "p"):
        out[f"{f}_20"] = getattr(st, f).numpy().copy()
    out["theta_20"] = s.theta.numpy().copy()
    out["keys"] = np.array(DIAG_KEYS)
    return out


def pid_case(steps=30, target=0.8):
    spec = CASES["box"]
    fx = dict(np.load(OUT / "ns3d_thermal_box.npz"))
    bodies = [RMesh(fx[f"mesh{i}_v"], fx[f"mesh{i}_f"], name=f"b{i}", thermal=m, thermal_value=val)
              for i, (_, _, _, m, val) in enumerate(spec["bodies"])]
    cfg = RConfig(**spec["grid"], bodies=bodies, device="cpu", dtype="float64", dye=False, perturb=False,
                  ref_area=0.3, length_ref=0.8, span_ref=0.6, moment_ref=(0.2, 0.0, 0.5),
                  thermal=RThermal(**spec["thermal"]), heat_sources=[RHeat(**h) for h in spec["heaters"]],
                  adaptive_dt="pid", pid_target=target)
    ops = rcore.make_operators3d(cfg)
    g = rgeo.build_geometry3d(bodies, ops, use_shared=False)
    phys = Physics3D(cfg, ops, g, bodies)
    z = torch.zeros((), dtype=torch.float64)
    st = rcore.State3D(torch.as_tensor(fx["u0"]), torch.as_tensor(fx["v0"]), torch.as_tensor(fx["w0"]),
                       torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z.clone(), z.clone())
    theta = torch.as_tensor(fx["theta0"])
    dts, mem = [], []
    with torch.no_grad():
        for _ in range(steps):
            r = rcore.step3d_phys(st, g, ops, phys=phys, theta=theta)
            st, theta = r.state, r.theta
            dts.append(float(st.dt))
            mem.append([float(phys.pid.dt_prev), float(phys.pid.e1), float(phys.pid.e2)])
    out = {"dt": np.array(dts), "ctrl": np.array(mem), "pid_target": np.float64(target), "steps": np.int64(steps)}
    for f in ("u", "v", "w", "p"):
        out[f"{f}_{steps}"] = getattr(st, f).numpy().copy()
    out[f"theta_{steps}"] = theta.numpy().copy()
    return out


def main():
    torch.set_num_threads(8)
    if sys.argv[1:] == ["pid"]:
        np.savez_compressed(OUT / "ns3d_pid_box.npz", **pid_case())
        print("wrote pid_box", flush=True)
        return
    if sys.argv[1:] == ["diag"]:
        np.savez_compressed(OUT / "ns3d_diag_box.npz", **diag_case())
        print("wrote diag_box", flush=True)
        return
    np.savez_compressed(OUT / "ns3d_pid_box.npz", **pid_case())
    for name, spec in CASES.items():
        np.savez_compressed(OUT / f"ns3d_thermal_{name}.npz", **run(name, spec))
        print("wrote", name, flush=True)
    np.savez_compressed(OUT / "ns3d_diag_box.npz", **diag_case())
    print("wrote diag_box", flush=True)


if __name__ == "__main__":
    main()
