"""Write the 3D golden fixtures from the reference solver (run in the build container only).

The reference is imported from ``/root/reference/src`` (read-only) and run on the CPU. The
fixtures pin the composition of whole 3D steps (stage order, inlet/outlet correction before
penalisation, periodic sync after it, step-averaged pressure, the six coefficients with and
without a symmetry plane) for every boundary family and closure, and the reference's mask
builder on triangulated boxes.

Each file holds inputs and expected outputs only (numpy arrays, no code):

* ``ns3d_<case>_f64.npz`` for the small cases of ``tests/cases.py`` (24 x 16 x 16): the body
  masks (built here by this package's builder and fed to both sides), the seeded start
  (noise, boundary values, projection by the reference), u, v, w, p after steps 1 and 5, dt,
  t and the coefficients of every step.
* ``ns3d_wale_moving_py_f32.npz``: the car miniature in float32, 20 steps.
* ``ns3d_masks.npz``: the reference's ``build_geometry3d`` on two overlapping triangulated
  boxes, one spinning, on the 24 x 16 x 16 grid (float64).
* ``ns3d_dye_<case>.npz``: the reference's smoke on the laminar tunnel case and the periodic-y
  car case: the automatic rake (positions, nozzle window and weights) and six steps of MUSCL
  advection with the step-5 velocity, solids and (pulsed) nozzle sources (float64).

Usage: ``python tests/golden/make_ns3d_fixtures.py`` (under a minute on 8 cores).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent
ROOT = OUT.parent.parent
sys.path.insert(0, str(REF))
sys.path.insert(0, str(ROOT))

import torch  # noqa: E402

from aerolume.solvers.ns3d import core as rcore  # noqa: E402
from aerolume.solvers.ns3d.config import ImmersedMesh as RMesh  # noqa: E402
from aerolume.solvers.ns3d.config import SolverConfig3D as RConfig  # noqa: E402
from aerolume.solvers.ns3d.geometry import Geometry3D as RGeometry  # noqa: E402
from aerolume.solvers.ns3d.geometry import build_geometry3d as rbuild  # noqa: E402

from aerolume_amd.ns3d import setup as su  # noqa: E402
from tests.cases import CASES3D, GRID3D, SEED3D, box_mesh, cfg3d  # noqa: E402

GEOM_KEYS = ("chi_u", "chi_v", "chi_w", "chi_c", "a_u", "a_v", "a_w", "k_u", "k_v", "k_w")


def ref_config(name, dtype):
    kw = dict(GRID3D, **CASES3D[name])
    return RConfig(**kw, bodies=[], dtype=dtype, device="cpu", dye=False, perturb=False, ref_area=0.3,
                   length_ref=0.8, span_ref=0.6, moment_ref=(0.2, 0.0, 0.5))


def ref_geometry(name, dtype):
    cfg = cfg3d(name, "float64")
    ops = su.make_operators3d(cfg, device="cpu")
    g = su.build_geometry3d(cfg.bodies, ops)
    t = torch.float64 if dtype == "float64" else torch.float32
    if g.empty:
        z = torch.zeros(0, 0, 0, dtype=t)
        k = torch.zeros(0, 3, 0, 0, 0, dtype=t)
        return g, RGeometry((0, 0, 0, 0, 0, 0), z, z, z, z, z, z, z, k, k, k, g.nbody)
    return g, RGeometry(g.window, *(getattr(g, k).to(t) for k in GEOM_KEYS), g.nbody)


def noisy_start(ops, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    nu_ = rng.standard_normal((nz, ny, nx + 1))
    nv_ = rng.standard_normal((nz, ny + 1, nx))
    nw_ = rng.standard_normal((nz + 1, ny, nx))
    base = 0.0 if (all(ops.periodic_axes) or ops.closed) else 1.0
    u = torch.as_tensor(base + 0.01 * nu_, dtype=ops.dtype)
    v = torch.as_tensor(0.01 * nv_, dtype=ops.dtype)
    w = torch.as_tensor(0.01 * nw_, dtype=ops.dtype)
    u, v, w = rcore.apply_boundary_values(u, v, w, ops)
    u, v, w, _ = rcore.project3d(u, v, w, ops)
    return u, v, w


def run_case(name, dtype, steps, keep):
    ops = rcore.make_operators3d(ref_config(name, dtype))
    g, rg = ref_geometry(name, dtype)
    u, v, w = noisy_start(ops, SEED3D)
    z = torch.zeros((), dtype=ops.dtype)
    st = rcore.State3D(u, v, w, torch.zeros(ops.nz, ops.ny, ops.nx, dtype=ops.dtype), z.clone(), z.clone())
    out = {"u0": u.numpy(), "v0": v.numpy(), "w0": w.numpy(), "g_window": np.asarray(g.window, np.int64),
           "g_nbody": np.int64(g.nbody)}
    if not g.empty:
        for k in GEOM_KEYS:
            out["g_" + k] = getattr(g, k).numpy()
    dts, ts, coeffs = [], [], []
    with torch.no_grad():
        for n in range(1, steps + 1):
            st, c = rcore.step3d_full(st, rg, ops)
            dts.append(float(st.dt))
            ts.append(float(st.t))
            coeffs.append(c.numpy().copy())
            if n in keep:
                for f in ("u", "v", "w", "p"):
                    out[f"{f}_{n}"] = getattr(st, f).numpy().copy()
    out["dt"] = np.array(dts)
    out["t"] = np.array(ts)
    out["coeffs"] = np.stack(coeffs)
    return out


MASK_BOXES = (((0.2, 0.0, 0.9), (0.8, 0.6, 0.5), 0.0, (0.0, 0.0, 0.0)),
              ((0.7, 0.15, 0.8), (0.5, 0.4, 0.6), -1.5, (0.7, 0.15, 0.8)))


def masks():
    rc = ref_config("laminar", "float64")
    ops = rcore.make_operators3d(rc)
    bodies = []
    out = {}
    for i, (c, s, om, cen) in enumerate(MASK_BOXES):
        v, f = box_mesh(c, s)
        bodies.append(RMesh(v, f, omega=om, center=cen, axis=(0.0, 1.0, 0.0)))
        out[f"center_{i}"], out[f"size_{i}"] = np.array(c), np.array(s)
        out[f"omega_{i}"], out[f"rot_center_{i}"] = np.float64(om), np.array(cen)
    g = rbuild(bodies, ops, use_shared=False)
    out["window"] = np.asarray(g.window, np.int64)
    for k in GEOM_KEYS:
        out[k] = getattr(g, k).numpy()
    return out


def dye_case(name, pulse):
    from aerolume.solvers.ns3d import dye as rdye
    ops = rcore.make_operators3d(ref_config(name, "float64"))
    g, rg = ref_geometry(name, "float64")
    fx = np.load(OUT / f"ns3d_{name}_f64.npz")
    cfg = cfg3d(name, "float64")
    pts = np.vstack([b.vertices_numpy() for b in cfg.bodies])
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    bounds = (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    extent = (ops.x0, ops.x0 + ops.nx * ops.h, ops.y0, ops.y0 + ops.ny * ops.h, ops.z0, ops.z0 + ops.nz * ops.h)
    rake = rdye.default_dye_rake(bounds, extent, ops.h, kind="rake", symmetry=ops.symmetry,
                                 ground=ops.bc_zmin in ("noslip", "moving"))
    if pulse:
        rake.pulse_on, rake.pulse_off = 0.05, 0.03
    src = rdye.build_dye_sources(rake, ops)
    u, v, w = (torch.as_tensor(fx[f"{f}_5"]) for f in "uvw")
    dt = torch.tensor(float(fx["dt"][4]), dtype=torch.float64)
    t = torch.tensor(float(fx["t"][4]), dtype=torch.float64)
    c = torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64)
    out = {"bounds": np.array(bounds), "positions": rake.positions, "radius": np.float64(rake.radius),
           "pulse": np.array([rake.pulse_on, rake.pulse_off]), "window": np.asarray(src.window, np.int64),
           "weight": src.weight.numpy(), "dt": dt.numpy(), "t0": t.numpy()}
    for n in range(1, 7):
        c = rdye.advect_dye3d(c, u, v, w, dt, ops)
        t = t + dt
        rdye.apply_solids_and_sources(c, t, rg, src)
        out[f"c_{n}"] = c.numpy().copy()
    return out


def main():
    torch.set_num_threads(8)
    for name in CASES3D:
        np.savez_compressed(OUT / f"ns3d_{name}_f64.npz", **run_case(name, "float64", 5, (1, 5)))
        print("wrote", name, flush=True)
    np.savez_compressed(OUT / "ns3d_wale_moving_py_f32.npz", **run_case("wale_moving_py", "float32", 20, (20,)))
    np.savez_compressed(OUT / "ns3d_masks.npz", **masks())
    np.savez_compressed(OUT / "ns3d_dye_laminar.npz", **dye_case("laminar", pulse=False))
    np.savez_compressed(OUT / "ns3d_dye_wale_moving_py.npz", **dye_case("wale_moving_py", pulse=True))
    print("done")


if __name__ == "__main__":
    main()
