"""Write the 3D traced-step golden fixtures from the reference solver (build container only).

The reference is imported from ``/root/reference/src`` (read-only) and run on the CPU in
float64. ``core.step3d_phys(..., trace=records)`` fills one record per RK stage (the explicit
tendency, the stage value before penalisation, after penalisation and buoyancy, the
projection's right-hand side, psi, the projected velocity and, with heat transfer, theta);
``FlowSolver3D.trace`` / ``trace_step`` build their outputs from exactly these records.

Each file holds inputs and expected outputs only (numpy arrays):

* ``ns3d_trace_wale_moving_py.npz``: the car miniature of ``tests/cases.py`` (WALE, moving
  floor, periodic y; masks from this package's builder fed to the reference, as the 3D step
  fixtures do), traced from the reference's seeded start stored in
  ``ns3d_wale_moving_py_f64.npz``.
* ``ns3d_trace_box.npz``: the heated box case (``core.step3d_phys`` with a ``Physics3D``
  carrying heat transfer), traced from the start stored in ``ns3d_thermal_box.npz``.

Keys: ``dt``, ``p`` (the step's pressure sum_s w_s psi_s / dt), and per stage s
``s<s>_<field>`` with field in tend_u|v|w, star_u|v|w, pen_u|v|w, rhs, psi,
proj_u|v|w (theta with heat transfer), staggered as the state.

Usage: ``python tests/golden/make_ns3d_trace_fixtures.py`` (seconds).
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

from tests.golden import make_ns3d_fixtures as m3  # noqa: E402
from tests.golden import make_ns3d_thermal_fixtures as mt  # noqa: E402

RECORD = (("tendency", "tend"), ("u_star", "star"), ("u_penalised", "pen"), ("u_projected", "proj"))


def flatten(records, new, dt):
    out = {"dt": np.float64(dt), "p": new.p.numpy().copy()}
    for r in records:
        s = r["stage"]
        for key, tag in RECORD:
            for c, x in zip("uvw", r[key]):
                out[f"s{s}_{tag}_{c}"] = x.numpy().copy()
        out[f"s{s}_rhs"] = r["rhs"].numpy().copy()
        out[f"s{s}_psi"] = r["psi"].numpy().copy()
        if "theta" in r:
            out[f"s{s}_theta"] = r["theta"].numpy().copy()
    return out


def car():
    name = "wale_moving_py"
    fx = np.load(OUT / f"ns3d_{name}_f64.npz")
    ops = rcore.make_operators3d(m3.ref_config(name, "float64"))
    _, rg = m3.ref_geometry(name, "float64")
    z = torch.zeros((), dtype=torch.float64)
    st = rcore.State3D(*(torch.as_tensor(fx[k]) for k in ("u0", "v0", "w0")),
                       torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z.clone(), z.clone())
    rec: list = []
    with torch.no_grad():
        r = rcore.step3d_phys(st, rg, ops, trace=rec)
    assert np.allclose(r.state.u.numpy(), fx["u_1"], rtol=0, atol=1e-13)
    return flatten(rec, r.state, r.state.dt)


def box():
    spec = mt.CASES["box"]
    fx = np.load(OUT / "ns3d_thermal_box.npz")
    bodies = [mt.RMesh(v, f, name=f"b{i}", thermal=m, thermal_value=val)
              for i, (v, f, m, val) in enumerate(mt.meshes(spec))]
    cfg = mt.RConfig(**spec["grid"], bodies=bodies, device="cpu", dtype="float64", dye=False, perturb=False,
                     ref_area=0.3, length_ref=0.8, span_ref=0.6, moment_ref=(0.2, 0.0, 0.5),
                     thermal=mt.RThermal(**spec["thermal"]), heat_sources=[mt.RHeat(**h) for h in spec["heaters"]])
    ops = rcore.make_operators3d(cfg)
    g = mt.rgeo.build_geometry3d(bodies, ops, use_shared=False)
    phys = mt.Physics3D(cfg, ops, g, bodies)
    z = torch.zeros((), dtype=torch.float64)
    st = rcore.State3D(*(torch.as_tensor(fx[k]) for k in ("u0", "v0", "w0")),
                       torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z.clone(), z.clone())
    rec: list = []
    with torch.no_grad():
        r = rcore.step3d_phys(st, g, ops, phys=phys, theta=torch.as_tensor(fx["theta0"]), trace=rec)
    assert np.allclose(r.theta.numpy(), fx["theta_1"], rtol=0, atol=1e-13)
    return flatten(rec, r.state, r.state.dt)


def main():
    torch.set_num_threads(8)
    np.savez_compressed(OUT / "ns3d_trace_wale_moving_py.npz", **car())
    np.savez_compressed(OUT / "ns3d_trace_box.npz", **box())
    print("done")


if __name__ == "__main__":
    main()
