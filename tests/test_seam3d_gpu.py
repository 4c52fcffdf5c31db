"""GPU checks of the 3D kernel seam: ``seam.fast_step`` / ``fast_step_phys`` called as the
reference's ``_step_fn`` calls them, with ops, geometry and state objects that carry exactly
the reference's attributes (``tests/golden/ref_interface_fields.json``) and a state of the
reference's NamedTuple shape. Results equal this package's functional steps bit for bit and
follow the reference's trajectories (float64, 1e-10 relative)."""
from __future__ import annotations

import collections
import json
import types
from pathlib import Path

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _fixture(name):
    path = GOLDEN / name
    if not path.exists():
        pytest.skip(f"missing fixture {name}")
    return dict(np.load(path)) if name.endswith(".npz") else json.loads(path.read_text())


def _like_reference(obj, spec):
    """A plain object with only the reference type's fields and properties."""
    ns = types.SimpleNamespace(**{f: getattr(obj, f) for f in spec["fields"] if hasattr(obj, f)})
    for m in spec["methods"]:
        if m == "empty":
            ns.empty = obj.empty
    return ns


def _rel(a, b):
    a = a.detach().double().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1.0))


def test_fast_step_with_reference_shaped_objects():
    _need_gpu()
    from aerolume_amd.ns3d import seam
    from aerolume_amd.ns3d import setup as su
    from aerolume_amd.ns3d import solver as S
    print("[seam gpu] fast_step", flush=True)
    spec = _fixture("ref_interface_fields.json")
    fx = _fixture("ns3d_wale_moving_py_f64.npz")
    cfg = cases.cfg3d("wale_moving_py", "float64", device="cuda")
    ops = S.make_operators3d(cfg)
    geom = su.build_geometry3d(cfg.bodies, ops)
    missing = [f for f in spec["Ops3D"]["fields"] if not hasattr(ops, f)]
    assert set(missing) <= {"poisson", "forcing"}, missing
    rops, rgeom = _like_reference(ops, spec["Ops3D"]), _like_reference(geom, spec["Geometry3D"])
    RefState = collections.namedtuple("State3D", spec["State3D"]["fields"])
    t = lambda a: torch.as_tensor(a, device="cuda", dtype=torch.float64)  # noqa: E731
    z = torch.zeros((), device="cuda", dtype=torch.float64)
    start = RefState(t(fx["u0"]), t(fx["v0"]), t(fx["w0"]), torch.zeros(ops.nz, ops.ny, ops.nx, device="cuda",
                                                                          dtype=torch.float64), z.clone(), z.clone())
    kern = seam.load_kernels(rops.advection, torch.device("cuda"), torch.float64)
    st = start
    for n in range(1, 6):
        st, c = seam.fast_step(st, rgeom, rops, kern)
        assert type(st) is RefState
        assert abs(float(st.dt) - fx["dt"][n - 1]) <= 1e-12 * fx["dt"][n - 1]
        assert np.abs(c.cpu().numpy() - fx["coeffs"][n - 1]).max() < 1e-9 * max(1.0, np.abs(fx["coeffs"][n - 1]).max())
    for f in "uvwp":
        assert _rel(getattr(st, f), fx[f"{f}_5"]) < 1e-10, f
    mine, _ = S.step3d_full(S.State3D(*start), geom, ops)
    ref1, _ = seam.fast_step(start, rgeom, rops, kern)
    for f in "uvwp":
        assert torch.equal(getattr(mine, f), getattr(ref1, f)), f


def test_fast_step_phys_with_reference_shaped_objects():
    _need_gpu()
    from aerolume_amd.ns3d import seam
    from aerolume_amd.ns3d import solver as S
    from aerolume_amd.ns3d import thermal as T
    print("[seam gpu] fast_step_phys", flush=True)
    spec = _fixture("ref_interface_fields.json")
    fx = _fixture("ns3d_thermal_box.npz")
    cfg = cases.thermal_cfg("box", fx, "float64", device="cuda")
    ops = S.make_operators3d(cfg)
    geom = S.build_geometry3d(cfg.bodies, ops)
    thermal = T.ThermalOps3D(cfg, ops, geom, cfg.bodies)
    rops, rgeom = _like_reference(ops, spec["Ops3D"]), _like_reference(geom, spec["Geometry3D"])
    RefState = collections.namedtuple("State3D", spec["State3D"]["fields"])
    t = lambda a: torch.as_tensor(np.asarray(a), device="cuda", dtype=torch.float64)  # noqa: E731
    z = torch.zeros((), device="cuda", dtype=torch.float64)
    st = RefState(t(fx["u0"]), t(fx["v0"]), t(fx["w0"]), torch.zeros(ops.nz, ops.ny, ops.nx, device="cuda",
                                                                       dtype=torch.float64), z.clone(), z.clone())
    theta = t(fx["theta0"])
    kern = seam.load_kernels(rops.advection, torch.device("cuda"), torch.float64)
    phys = types.SimpleNamespace(thermal=thermal, fans=None, ports=None, paint=None)
    out = seam.fast_step_phys(st, rgeom, rops, kern, phys=phys, theta=theta)
    assert type(out.state) is RefState
    assert _rel(out.theta, fx["theta_1"]) < 1e-10
    for f in "uvw":
        assert _rel(getattr(out.state, f), fx[f"{f}_1"]) < 1e-10, f
    assert np.abs(thermal.heat_rate.cpu().numpy() - fx["heat_rate"][0]).max() < 1e-9 * max(1.0, np.abs(fx["heat_rate"][0]).max())
    with pytest.raises(NotImplementedError):
        seam.fast_step_phys(st, rgeom, rops, kern, phys=types.SimpleNamespace(thermal=thermal, fans=[1]), theta=theta)
