"""3D heat transfer on the MI355X through the C ABI: the HIP step against the reference's own
trajectories (float64, ``tests/golden/ns3d_thermal_*.npz``) and the oracle (float32), graph
replays against eager steps, and the conservation properties of the reference's tests."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle.ns3d as O
from aerolume_amd.ns3d import setup as su
from aerolume_amd.ns3d import solver as S3
from aerolume_amd.ns3d import thermal as T
from aerolume_amd.ns3d.config import HeatSource3D, SolverConfig3D, ThermalConfig3D
from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = list(cases.THERMAL3D)


def _fixture(name):
    path = GOLDEN / f"ns3d_thermal_{name}.npz"
    if not path.exists():
        pytest.skip(f"missing fixture {path.name}")
    return dict(np.load(path))


def _gpu_setup(name, fx, dtype="float64"):
    cfg = cases.thermal_cfg(name, fx, dtype, device="cuda")
    ops = S3.make_operators3d(cfg)
    g = S3.build_geometry3d(cfg.bodies, ops)
    th = T.ThermalOps3D(cfg, ops, g, cfg.bodies)
    return cfg, ops, g, th


def _state(fx, ops):
    t = lambda a: torch.as_tensor(np.asarray(a), device=ops.device, dtype=ops.dtype).contiguous()  # noqa: E731
    z = torch.zeros((), device=ops.device, dtype=ops.dtype)
    return S3.State3D(t(fx["u0"]), t(fx["v0"]), t(fx["w0"]), torch.zeros(ops.nz, ops.ny, ops.nx, device=ops.device,
                                                                           dtype=ops.dtype), z.clone(), z.clone()), t(fx["theta0"])


def _rel(a, b):
    a = a.detach().double().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", NAMES)
def test_thermal_steps_match_reference_float64(name):
    print(f"[thermal gpu] reference case {name} float64", flush=True)
    fx = _fixture(name)
    cfg, ops, g, th = _gpu_setup(name, fx)
    st, theta = _state(fx, ops)
    steps = int(fx["steps"])
    for n in range(1, steps + 1):
        out = S3.step3d_phys(st, g, ops, thermal=th, theta=theta)
        st, theta = out.state, out.theta
        assert float(st.dt) == pytest.approx(float(fx["dt"][n - 1]), rel=1e-12)
        assert np.allclose(out.coeffs.double().cpu().numpy(), fx["coeffs"][n - 1], rtol=1e-9, atol=1e-10)
        assert np.allclose(th.heat_rate.double().cpu().numpy(), fx["heat_rate"][n - 1], rtol=1e-9, atol=1e-10)
        if n in (1, steps):
            for f in ("u", "v", "w", "p"):
                assert _rel(getattr(st, f), fx[f"{f}_{n}"]) < 1e-10, (f, n)
            err = float(np.abs(theta.double().cpu().numpy() - fx[f"theta_{n}"]).max())
            assert err < 1e-11, ("theta", n, err)
    walls = th.wall_heat_rates(theta, st.u)
    assert np.allclose([walls[s] for s in O.WALL_SIDES], fx["wall_heat"], rtol=1e-9, atol=1e-10)
    assert np.allclose(th.theta_wall.double().cpu().numpy(), fx["wall_theta"], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("name", ["box", "tunnel_les"])
def test_thermal_float32_matches_oracle(name):
    print(f"[thermal gpu] oracle case {name} float32", flush=True)
    fx = _fixture(name)
    cfg, ops, g, th = _gpu_setup(name, fx, "float32")
    st, theta = _state(fx, ops)
    ccfg = cases.thermal_cfg(name, fx, "float32")
    cops = su.make_operators3d(ccfg, device="cpu")
    cg = su.build_geometry3d(ccfg.bodies, cops)
    oth = O.Thermal(T.build_thermal(ccfg, cops, ccfg.bodies, cg.window), cops, ccfg.re)
    z = torch.zeros((), dtype=torch.float32)
    ost = O.State3D(*(torch.as_tensor(fx[k], dtype=torch.float32) for k in ("u0", "v0", "w0")),
                    torch.zeros(cops.nz, cops.ny, cops.nx), z.clone(), z.clone())
    otheta = torch.as_tensor(fx["theta0"], dtype=torch.float32)
    for n in range(5):
        out = S3.step3d_phys(st, g, ops, thermal=th, theta=theta)
        st, theta = out.state, out.theta
        ost, oc, otheta = O.step3d_phys(ost, cg, cops, oth, otheta)
    # float32 tolerances: five steps of the same arithmetic in a different summation order
    assert float(st.dt) == pytest.approx(float(ost.dt), rel=1e-5)
    for f in ("u", "v", "w"):
        assert float((getattr(st, f).cpu() - getattr(ost, f)).abs().max()) < 2e-5, f
    assert float((theta.cpu() - otheta).abs().max()) < 2e-5
    assert np.allclose(th.heat_rate.cpu().numpy(), oth.heat_rate.numpy(), rtol=1e-3, atol=1e-4)


def _box(n=12, **kw):
    base = dict(nx=n, ny=n, nz=n, h=1.0 / n, x0=0.0, y0=0.0, z0=0.0, re=20.0, domain="closed", device="cuda",
                dtype="float64", dye=False, perturb=False)
    base.update(kw)
    return SolverConfig3D(**base)


def test_solver_graphs_replay_eager_steps_bit_for_bit():
    print("[thermal gpu] graphs vs eager", flush=True)
    fx = _fixture("box")
    runs = []
    for graphs in (True, False):
        cfg = cases.thermal_cfg("box", fx, "float32", device="cuda", use_cuda_graphs=graphs)
        s = S3.FlowSolver3D(cfg)
        s.set_velocity(fx["u0"], fx["v0"], fx["w0"])
        s.set_temperature(fx["theta0"])
        s.advance(23)
        runs.append(s)
    a, b = runs
    assert torch.equal(a.theta, b.theta)
    assert torch.equal(a.state.w, b.state.w)
    ha, hb = a.physics_history(), b.physics_history()
    assert len(ha["t"]) == 23 and np.array_equal(ha["heat_rate"], hb["heat_rate"])
    assert np.array_equal(ha["theta_wall_1"], hb["theta_wall_1"])
    q = a.heat_rates()
    assert q[0]["heat_rate"] == pytest.approx(ha["heat_rate_0"][-1], rel=1e-6)
    assert q[2]["heat_rate"] == 0.0 and math.isnan(q[2]["nusselt"])


def test_solver_conduction_between_walls_is_exact():
    n = 12
    x = (np.arange(n) + 0.5) / n
    cfg = _box(n, ny=4, nz=4, re=10.0, thermal=ThermalConfig3D(
        prandtl=1.0, walls={"left": ("temperature", 0.5), "right": ("temperature", -0.5)}))
    s = S3.FlowSolver3D(cfg)
    s.set_temperature(np.broadcast_to(0.5 - x, (4, 4, n)).copy())
    s.advance(20)
    th = s.field("temperature").double().cpu().numpy()
    assert np.abs(th - (0.5 - x)[None, None, :]).max() < 1e-13
    w = s.wall_heat_rates()
    assert w["left"] == pytest.approx((4 / n) ** 2, rel=1e-12)
    assert float(s.field("speed").max()) < 1e-14


def test_solver_heat_source_raises_the_stored_heat_at_its_power():
    cfg = _box(thermal=ThermalConfig3D(prandtl=0.71, richardson=0.5),
               heat_sources=[HeatSource3D(kind="sphere", center=(0.4, 0.5, 0.55), size=(0.2,), power=3.0),
                             HeatSource3D(kind="box", center=(0.7, 0.3, 0.3), size=(0.2, 0.1, 0.3), power=-1.0)])
    s = S3.FlowSolver3D(cfg)
    s.advance(25)
    stored = float(s.theta.double().sum()) * s.ops.h ** 3 * s.thermal.pe
    assert stored / s.t == pytest.approx(2.0, rel=1e-10)


def test_solver_stratified_fluid_stays_at_rest():
    n = 10
    cfg = _box(n, bc_zmin="noslip", bc_zmax="noslip", thermal=ThermalConfig3D(
        prandtl=0.71, richardson=5.0, walls={"bottom": ("temperature", -0.5), "top": ("temperature", 0.5)}))
    s = S3.FlowSolver3D(cfg)
    z = (np.arange(n) + 0.5) / n - 0.5
    s.set_temperature(np.broadcast_to(z[:, None, None], (n, n, n)).copy())
    s.advance(60)
    assert float(s.field("speed").max()) < 1e-12
    assert np.abs(s.field("temperature").cpu().numpy()[:, 0, 0] - z).max() < 1e-12


def test_live_reynolds_and_wall_values_keep_graphs():
    fx = _fixture("box")
    cfg = cases.thermal_cfg("box", fx, "float64", device="cuda")
    s = S3.FlowSolver3D(cfg)
    s.advance(10)
    pe0 = s.thermal.pe
    s.set_inflow(re=160.0)
    assert s.thermal.pe == pytest.approx(2 * pe0)
    assert float(s.thermal.scalars[0]) == pytest.approx(1.0 / (160.0 * 0.71), rel=1e-15)
    s.set_thermal(walls={"bottom": ("heat_flux", 1.4), "left": ("temperature", 0.5)})
    assert float(s.thermal.side_vals[4]) == pytest.approx(1.4)
    s.advance(10)
    assert bool(torch.isfinite(s.theta).all())
    s.set_thermal(richardson=0.0)
    s.advance(5)
    assert s.thermal.ri == 0.0 and bool(torch.isfinite(s.theta).all())


def test_solver_diagnostics_match_reference_rows(tmp_path):
    """Diagnostics rows of the HIP step (steps 10 and 20 of a 7 + 13 step advance) against the
    reference FlowSolver3D's rows from the same start (float64); round-off residuals by
    magnitude; graph and eager runs give identical rows."""
    print("[thermal gpu] diagnostics rows", flush=True)
    fx = _fixture("box")
    ref = dict(np.load(GOLDEN / "ns3d_diag_box.npz"))
    keys = [str(k) for k in ref["keys"]]
    assert tuple(keys) == S3.DIAG_KEYS
    rows = []
    for graphs in (True, False):
        cfg = cases.thermal_cfg("box", fx, "float64", device="cuda", diagnostics=True, use_cuda_graphs=graphs)
        s = S3.FlowSolver3D(cfg)
        s.set_velocity(fx["u0"], fx["v0"], fx["w0"])
        s.set_temperature(fx["theta0"])
        s.advance(7)
        s.advance(13)
        d = s.diagnostics()
        assert list(d["step"]) == [10.0, 20.0]
        rows.append(np.stack([d[k] for k in keys], axis=1))
        for k in keys:
            if k in ("div_rms", "div_max", "poisson_residual"):
                assert np.all(d[k] < 1e-12) and np.all(ref[k] < 1e-12), k
            else:
                assert np.allclose(d[k], ref[k], rtol=1e-9, atol=1e-13), (k, d[k], ref[k])
        assert float(np.abs(s.theta.double().cpu().numpy() - ref["theta_20"]).max()) < 1e-11
    assert np.array_equal(rows[0], rows[1])
    path = s.diagnostics_csv(tmp_path / "diag.csv")
    text = open(path).read().splitlines()
    assert text[0].startswith("step,t,dt,cfl") and len(text) == 3
    assert s.latest_diagnostics()["step"] == 20.0


# ---------------------------------------------------------------------------
# PID step (adaptive_dt="pid")
# ---------------------------------------------------------------------------

def _pid_fixture():
    path = GOLDEN / "ns3d_pid_box.npz"
    if not path.exists():
        pytest.skip("missing fixture ns3d_pid_box.npz")
    return dict(np.load(path))


def test_pid_steps_match_reference_float64():
    """dt and the controller memory of each of the reference's 30 PID steps on the heated box
    (pid_target 0.8), and the fields after them, through the functional step."""
    print("[thermal gpu] PID steps float64 vs reference", flush=True)
    fx, ref = _fixture("box"), _pid_fixture()
    cfg, ops, g, th = _gpu_setup("box", fx)
    st, theta = _state(fx, ops)
    pid = S3.PidDt3(ops, float(ref["pid_target"]))
    n = int(ref["steps"])
    for k in range(n):
        out = S3.step3d_phys(st, g, ops, thermal=th, theta=theta, pid=pid)
        st, theta = out.state, out.theta
        assert float(st.dt) == pytest.approx(float(ref["dt"][k]), rel=1e-12), k
        assert np.allclose(pid.mem.cpu().numpy(), ref["ctrl"][k], rtol=1e-10, atol=1e-13), k
    for f in ("u", "v", "w", "p"):
        assert _rel(getattr(st, f), ref[f"{f}_{n}"]) < 1e-10, f
    assert float(np.abs(theta.double().cpu().numpy() - ref[f"theta_{n}"]).max()) < 1e-11


def test_pid_solver_graphs_chained_and_float32():
    """The FlowSolver3D PID run (graphs on and off, 7 + 23 steps, chained dt) equals the
    reference's 30 steps; reset starts the controller over; float32 against the oracle."""
    print("[thermal gpu] PID FlowSolver3D", flush=True)
    fx, ref = _fixture("box"), _pid_fixture()
    n = int(ref["steps"])
    runs = []
    for graphs in (True, False):
        cfg = cases.thermal_cfg("box", fx, "float64", device="cuda", use_cuda_graphs=graphs, adaptive_dt="pid",
                                pid_target=float(ref["pid_target"]))
        s = S3.FlowSolver3D(cfg)
        s.set_velocity(fx["u0"], fx["v0"], fx["w0"])
        s.set_temperature(fx["theta0"])
        s.advance(7)
        s.advance(n - 7)
        h = s.history()
        assert np.allclose(np.diff(np.r_[0.0, h["t"]]), ref["dt"], rtol=1e-11, atol=0)
        assert np.allclose(s._pid.mem.cpu().numpy(), ref["ctrl"][-1], rtol=1e-10, atol=1e-13)
        assert _rel(s.state.u, ref[f"u_{n}"]) < 1e-10
        runs.append(s)
    a, b = runs
    for f in ("u", "v", "w", "p", "t", "dt"):
        assert torch.equal(getattr(a.state, f), getattr(b.state, f)), f
    assert torch.equal(a._pid.mem, b._pid.mem)
    sd = a.state_dict()
    assert "pid" in sd
    a.reset()
    assert float(a._pid.mem.abs().max()) == 0.0
    # float32 against the oracle over 8 steps
    cfg, ops, g, th = _gpu_setup("box", fx, "float32")
    st, theta = _state(fx, ops)
    pid = S3.PidDt3(ops, float(ref["pid_target"]))
    ocfg = cases.thermal_cfg("box", fx, "float32")
    oops = su.make_operators3d(ocfg, device="cpu")
    og = su.build_geometry3d(ocfg.bodies, oops)
    oth = O.Thermal(T.build_thermal(ocfg, oops, ocfg.bodies, og.window), oops, ocfg.re)
    opid = O.PidDt3(float(ref["pid_target"]), dtype=torch.float32)
    z = torch.zeros((), dtype=torch.float32)
    f32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32)  # noqa: E731
    ost = O.State3D(f32(fx["u0"]), f32(fx["v0"]), f32(fx["w0"]), torch.zeros(oops.nz, oops.ny, oops.nx), z.clone(),
                    z.clone())
    otheta = f32(fx["theta0"])
    for _ in range(8):
        out = S3.step3d_phys(st, g, ops, thermal=th, theta=theta, pid=pid)
        st, theta = out.state, out.theta
        ost, _, otheta = O.step3d_phys(ost, og, oops, oth, otheta, pid=opid)
        assert float(st.dt) == pytest.approx(float(ost.dt), rel=2e-5)
    assert float((st.u.cpu() - ost.u).abs().max()) < 2e-5
    assert float((theta.cpu() - otheta).abs().max()) < 2e-5
    # and its error against the float64 reference dt
    assert float(st.dt) == pytest.approx(float(ref["dt"][7]), rel=1e-4)


def _rms_close(mine, ref, tol=1e-12):
    """rms fields compare by their squares (the square root magnifies round-off near zero)."""
    return float(np.abs(mine ** 2 - ref ** 2).max()) <= tol * max(float((ref ** 2).max()), 1.0)


def test_solver_host_reports_match_reference():
    """Running means (time-weighted, sampled every 10 steps), slices, surface probes and the
    heat balance of the reference FlowSolver3D's 7 + 13 step run with averaging started after
    step 7 (float64), graphs on."""
    print("[thermal gpu] host reports", flush=True)
    fx = _fixture("box")
    ref = dict(np.load(GOLDEN / "ns3d_diag_box.npz"))
    if "averaging_time" not in ref:
        pytest.skip("fixture without host reports")
    cfg = cases.thermal_cfg("box", fx, "float64", device="cuda", diagnostics=True, use_cuda_graphs=True)
    s = S3.FlowSolver3D(cfg)
    s.set_velocity(fx["u0"], fx["v0"], fx["w0"])
    s.set_temperature(fx["theta0"])
    s.advance(7)
    s.start_averaging()
    s.advance(13)
    assert s.averaging_time == pytest.approx(float(ref["averaging_time"]), rel=1e-12)
    mf = s.mean_fields()
    assert set(mf) == {k[5:] for k in ref if k.startswith("mean_")}
    for k, v in mf.items():
        mine, r = v.double().cpu().numpy(), ref["mean_" + k]
        if k.endswith("_rms"):
            assert _rms_close(mine, r), k
        else:
            assert float(np.abs(mine - r).max()) <= 1e-9 * max(float(np.abs(r).max()), 1.0), k
    for key in [k for k in ref if k.startswith("slice_") or k.startswith("mslice_")]:
        pre, axis, name = key.split("_", 2)
        pos = {"x": 0.37, "y": 0.5, "z": 0.21}[axis]
        mine = s.slice(axis, pos, name, mean=pre == "mslice").double().cpu().numpy()
        if name.endswith("_rms"):
            assert _rms_close(mine, ref[key]), key
        else:
            assert float(np.abs(mine - ref[key]).max()) <= 1e-9 * max(float(np.abs(ref[key]).max()), 1.0), key
    for pre, mean in (("surf_", False), ("msurf_", True)):
        rec = s.surface(mean=mean)[0]
        for k in ("vertices", "faces", "normals", "cp", "cf", "tau", "theta_wall", "heat_flux"):
            assert np.allclose(rec[k], ref[pre + k], rtol=1e-8, atol=1e-10), (pre, k)
    hb = s.heat_balance()
    assert list(hb) == [str(k) for k in ref["hb_keys"]]
    assert np.allclose([hb[k] for k in hb], ref["hb_values"], rtol=1e-8, atol=1e-10)
    t_last = s.history()["t"][-1]
    s.trim_history(0.1)
    assert s.history()["t"][0] >= t_last - 0.1 and list(s.diagnostics()["step"]) == [20.0]


def test_solver_set_bodies_equals_fresh_solver():
    """Bodies moved live: the next steps equal a solver built with the moved bodies and started
    from the same state (float64; masks, heat-equation data and plan rebuilt)."""
    print("[thermal gpu] set_bodies", flush=True)
    from aerolume_amd.ns3d.config import ImmersedMesh
    fx = _fixture("box")
    cfg = cases.thermal_cfg("box", fx, "float64", device="cuda", use_cuda_graphs=True)
    moved = [ImmersedMesh(b.vertices_numpy() + np.array([0.06, 0.0, 0.03]), b.faces, name=b.name,
                          thermal=b.thermal, thermal_value=b.thermal_value) for b in cfg.bodies]
    a = S3.FlowSolver3D(cfg)
    a.set_velocity(fx["u0"], fx["v0"], fx["w0"])
    a.set_temperature(fx["theta0"])
    a.advance(5)
    sd = a.state_dict()
    a.set_bodies(moved)
    a.advance(15)
    import dataclasses
    b = S3.FlowSolver3D(dataclasses.replace(cfg, bodies=moved))
    b.load_state_dict(sd)
    b.advance(15)
    for f in ("u", "v", "w", "p", "t", "dt"):
        assert torch.equal(getattr(a.state, f), getattr(b.state, f)), f
    assert torch.equal(a.theta, b.theta)
    assert np.array_equal(a.coefficients(), b.coefficients())
    assert len(a.history()["t"]) == 20


def test_solver_set_sources_equals_fresh_solver():
    """Heaters replaced live equal a solver configured with them from the same state (FP64)."""
    print("[thermal gpu] set_sources", flush=True)
    import dataclasses
    fx = _fixture("box")
    cfg = cases.thermal_cfg("box", fx, "float64", device="cuda", use_cuda_graphs=True)
    heaters = [HeatSource3D(**dict(h, power=2.0 * h.get("power", 1.0))) for h in cases.THERMAL3D["box"]["heaters"]]
    a = S3.FlowSolver3D(cfg)
    a.set_velocity(fx["u0"], fx["v0"], fx["w0"])
    a.set_temperature(fx["theta0"])
    a.advance(4)
    sd = a.state_dict()
    a.set_sources(heat_sources=heaters)
    a.advance(6)
    b = S3.FlowSolver3D(dataclasses.replace(cfg, heat_sources=heaters))
    b.load_state_dict(sd)
    b.advance(6)
    assert torch.equal(a.theta, b.theta) and torch.equal(a.state.w, b.state.w)
    with pytest.raises(NotImplementedError):
        a.set_sources(fans=[object()])
