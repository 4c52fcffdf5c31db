"""The 2D PID-controlled step (``adaptive_dt="pid"``) and the diagnostics rows on the MI355X
through the C ABI: the reference's trajectory and rows in float64
(``tests/golden/ns2d_pid_diag.npz``), the oracle in float32, and the FlowSolver's graphs and
chained steps against eager plain steps."""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle.ns2d as O
from aerolume_amd.ns2d import setup as su
from aerolume_amd.ns2d import solver as S
from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden" / "ns2d_pid_diag.npz"


def _fixture():
    if not GOLDEN.exists():
        pytest.skip("missing fixture ns2d_pid_diag.npz")
    return dict(np.load(GOLDEN))


def _rel(a, b):
    a = a.detach().double().cpu().numpy()
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-30)


def test_pid_steps_match_reference_float64():
    print("[pid gpu] float64 steps vs reference", flush=True)
    fx = _fixture()
    cfg = cases.pid_diag_cfg("float64", device="cuda")
    ops = S.make_operators(cfg)
    geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
    t = lambda a: torch.as_tensor(a, device=ops.device, dtype=ops.dtype)  # noqa: E731
    st = S.init_state(cfg, ops)._replace(u=t(fx["u0"]), v=t(fx["v0"]), theta=t(fx["theta0"]))
    assert st.ctrl is not None
    for n in range(30):
        st, _ = S.step_full(st, geom, ops)
        assert float(st.dt) == pytest.approx(float(fx["dt_hist"][n]), rel=1e-12), n
        assert np.allclose(st.ctrl.double().cpu().numpy(), fx["ctrl_hist"][n], rtol=1e-10, atol=1e-13), n
    for f in ("u", "v", "p", "theta"):
        assert _rel(getattr(st, f), fx[f"{f}30"]) < 1e-10, f


def test_pid_float32_matches_oracle():
    print("[pid gpu] float32 vs oracle", flush=True)
    fx = _fixture()
    res = []
    for dev in ("cuda", "cpu"):
        cfg = cases.pid_diag_cfg("float32", device=dev)
        ops = su.make_operators(cfg, device=dev)
        geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
        t = lambda a: torch.as_tensor(a, device=ops.device, dtype=ops.dtype)  # noqa: E731
        st = S.init_state(cfg, ops)._replace(u=t(fx["u0"]), v=t(fx["v0"]), theta=t(fx["theta0"]))
        step = S.step_full if dev == "cuda" else O.step_full
        with torch.no_grad():
            for _ in range(15):
                st, _ = step(st, geom, ops)
        res.append(st)
    g, c = res
    assert float(g.dt) == pytest.approx(float(c.dt), rel=2e-5)
    assert float(g.t) == pytest.approx(float(c.t), rel=2e-5)
    for f in ("u", "v", "theta"):
        assert float((getattr(g, f).cpu() - getattr(c, f)).abs().max()) < 5e-5, f


def test_pid_flowsolver_graphs_and_chained_equal_eager():
    print("[pid gpu] FlowSolver graphs / chained", flush=True)
    fx = _fixture()
    runs = []
    for graphs in (True, False):
        s = S.FlowSolver(dataclasses.replace(cases.pid_diag_cfg("float64", device="cuda"), use_cuda_graphs=graphs))
        s.set_velocity(fx["u0"], fx["v0"])
        s.set_temperature(fx["theta0"])
        s.advance(7)
        s.advance(13)
        runs.append(s)
    a, b = runs
    for f in ("u", "v", "p", "theta", "t", "dt", "ctrl"):
        assert torch.equal(getattr(a.state, f), getattr(b.state, f)), f
    assert _rel(a.state.u, fx["u20"]) < 1e-10 and _rel(a.state.theta, fx["theta20"]) < 1e-10
    assert float(a.state.dt) == pytest.approx(float(fx["dt20"]), rel=1e-12)
    assert np.allclose(a.state.ctrl.cpu().numpy(), fx["ctrl20"], rtol=1e-10, atol=1e-13)
    sd = a.state_dict()
    assert "ctrl" in sd and sd["ctrl"].numel() == 2


ROUNDOFF_COLS = ("div_rms", "div_max", "p_residual")


def _diag_solver(dtype, graphs, fx):
    cfg = dataclasses.replace(cases.pid_diag_cfg(dtype, device="cuda"), use_cuda_graphs=graphs, diagnostics=True)
    s = S.FlowSolver(cfg)
    s.set_velocity(fx["u0"], fx["v0"])
    s.set_temperature(fx["theta0"])
    s.advance(7)
    s.advance(13)
    return s


def test_diagnostics_rows_match_reference_float64(tmp_path):
    """Rows at steps 10 and 20 of a 7 + 13 step advance, graphs on and off, against the
    reference FlowSolver's rows; the continuity and Poisson residuals are round-off numbers
    of the direct solve and are compared by magnitude only."""
    print("[pid gpu] diagnostics rows float64 vs reference", flush=True)
    fx = _fixture()
    cols = [str(c) for c in fx["diag_columns"]]
    assert tuple(cols) == S.DIAG_COLUMNS
    runs = []
    for graphs in (True, False):
        s = _diag_solver("float64", graphs, fx)
        d = s.diagnostics()
        assert list(d["step"]) == [10.0, 20.0], graphs
        for c in cols[1:]:
            mine, theirs = d[c], fx[f"diag_{c}"]
            if c in ROUNDOFF_COLS:
                assert np.all(mine < 1e-13) and np.all(mine > 0), (c, mine)
            else:
                assert np.allclose(mine, theirs, rtol=1e-9, atol=1e-14), (c, mine, theirs)
        assert len(d["steps_per_s"]) == 2 and list(d["perf_steps"]) == [7.0, 20.0]
        runs.append(s)
    a, b = (r.diagnostics() for r in runs)
    for c in cols:
        assert np.array_equal(a[c], b[c]), c       # graph and eager rows bit for bit
    s = runs[0]
    last = s.latest_diagnostics()
    assert last["step"] == 20.0 and last["t"] == pytest.approx(float(fx["diag_t"][1]), rel=1e-12)
    path = s.diagnostics_csv(tmp_path / "diag.csv")
    lines = Path(path).read_text().splitlines()
    assert lines[0].split(",") == cols and len(lines) == 3
    assert float(lines[2].split(",")[0]) == 20.0
    # the rows stay with the run: a reset clears them, the next advance records afresh
    s.reset()
    assert len(s.diagnostics()["step"]) == 0
    s.advance(10)
    assert list(s.diagnostics()["step"]) == [10.0]


def test_diagnostics_rows_float32_vs_oracle_and_float64():
    """float32 rows against the oracle's float32 rows, and their error against the float64
    reference rows (velocity / temperature changes ~1e-6, pressure change ~1e-3 relative)."""
    print("[pid gpu] diagnostics rows float32 vs oracle", flush=True)
    from tests.test_oracle_ns2d_pid import oracle_diag_rows
    fx = _fixture()
    cols = [str(c) for c in fx["diag_columns"]][1:]
    d = _diag_solver("float32", True, fx).diagnostics()
    orc, _ = oracle_diag_rows(fx, dtype=torch.float32)
    for i, c in enumerate(cols):
        mine, o32, f64 = d[c], orc[:, i], fx[f"diag_{c}"]
        if c in ROUNDOFF_COLS:
            assert np.all(mine < 1e-5) and np.all(o32 < 1e-5), (c, mine, o32)
            continue
        tol = 5e-3 if c.startswith("dp") else 2e-5
        assert np.allclose(mine, o32, rtol=tol, atol=0), (c, mine, o32)
        assert np.allclose(mine, f64, rtol=tol, atol=0), (c, mine, f64)


def test_functional_diag_row_and_no_theta():
    """step_full(diag=True) returns the row of that one step (oracle, float64), with NaN
    temperature columns on a run without heat transfer."""
    print("[pid gpu] functional diag rows", flush=True)
    for thermal in (True, False):
        over = {} if thermal else dict(thermal=None, bodies=[], adaptive_dt="cfl")
        rows = []
        for dev in ("cuda", "cpu"):
            cfg = cases.pid_diag_cfg("float64", device=dev, **over)
            ops = su.make_operators(cfg, device=dev)
            geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
            st = S.init_state(cfg, ops)
            step = S.step_full if dev == "cuda" else O.step_full
            with torch.no_grad():
                st, _ = step(st, geom, ops)
                out = {}
                st, _ = step(st, geom, ops, out=out, diag=True)
            rows.append(out["diag"].double().cpu().numpy())
        g, c = rows
        assert g.shape == (22,)
        if not thermal:
            assert np.all(np.isnan(g[18:])) and np.all(np.isnan(c[18:]))
        n = 18 if not thermal else 22
        for i in range(n):
            if i in (3, 4, 5):
                assert g[i] < 1e-12 and c[i] < 1e-12, (i, g[i], c[i])
            else:
                assert g[i] == pytest.approx(c[i], rel=1e-9, abs=1e-14), (thermal, i, g[i], c[i])


def test_host_reports_match_reference():
    """surface(), heat_transfer(), heat_balance() (instantaneous and of the running means),
    mean_fields() and averaging_time after the reference's 7 + 13 step run with averaging
    started after step 7, against the reference FlowSolver's own reports (float64)."""
    print("[pid gpu] host reports vs reference", flush=True)
    fx = _fixture()
    cfg = dataclasses.replace(cases.pid_diag_cfg("float64", device="cuda"), use_cuda_graphs=True, diagnostics=True)
    s = S.FlowSolver(cfg)
    s.set_velocity(fx["u0"], fx["v0"])
    s.set_temperature(fx["theta0"])
    s.advance(7)
    s.start_averaging()
    s.advance(13)
    assert s.averaging_time == pytest.approx(float(fx["averaging_time"]), rel=1e-12)
    assert s.peclet == pytest.approx(400 * 0.71)
    for pre, mean in (("surf_", False), ("msurf_", True)):
        rec = s.surface(48, mean=mean)[0]
        for k in ("s", "x", "y", "nx", "ny", "cp", "cf", "theta_wall", "nu_local"):
            assert np.allclose(rec[k], fx[pre + k], rtol=1e-9, atol=1e-12), (pre, k)
    for pre, mean in (("ht_", False), ("mht_", True)):
        ht = s.heat_transfer(mean=mean)
        b = ht["bodies"][0]
        for k in ("perimeter", "heat_rate", "theta_wall", "nusselt"):
            assert b[k] == pytest.approx(float(fx[pre + k]), rel=1e-9), (pre, k)
        walls = [ht["walls"][w] for w in ("left", "right", "bottom", "top")]
        assert np.allclose(walls, fx[pre + "walls"], rtol=1e-9, atol=1e-13), pre
        assert ht["nusselt_mean"] == pytest.approx(float(fx[pre + "nusselt_mean"]), rel=1e-9)
    for pre, mean in (("hb_", False), ("mhb_", True)):
        hb = s.heat_balance(mean=mean)
        assert list(hb) == [str(k) for k in fx[pre + "keys"]]
        assert np.allclose([hb[k] for k in hb], fx[pre + "values"], rtol=1e-8, atol=1e-12), pre
    mf = s.mean_fields()
    for k in ("u", "v", "speed", "p", "cp", "vorticity", "u_rms", "v_rms", "p_rms", "cp_rms", "mask", "theta",
              "theta_rms"):
        ref, mine = fx["mean_" + k], mf[k].double().cpu().numpy()
        if k.endswith("_rms"):
            # sqrt(E[x^2] - E[x]^2): compare the variances, which carry the round-off linearly
            err = float(np.abs(mine ** 2 - ref ** 2).max())
            assert err <= 1e-13, (k, err)
            continue
        err = float(np.abs(mine - ref).max())
        assert err <= 1e-9 * max(float(np.abs(ref).max()), 1.0), (k, err)
    # trim_history keeps the rows of the last 0.2 convective times
    t_last = s.history()["t"][-1]
    s.trim_history(0.2)
    h = s.history()
    assert h["t"][0] >= t_last - 0.2 and h["t"][-1] == t_last
    d = s.diagnostics()
    assert list(d["step"]) == [20.0]


def test_set_bodies_keeps_flow_and_equals_fresh_solver():
    """Bodies replaced live: the next steps equal a solver built with the new bodies and started
    from the same state (the window moves, the plan is rebuilt, graphs re-captured)."""
    print("[pid gpu] set_bodies", flush=True)
    fx = _fixture()
    from aerolume_amd.ns2d.config import ImmersedBody
    new = [ImmersedBody(cases.circle(64, 0.6, 3.0, 1.8), thermal="temperature", thermal_value=1.0, name="moved")]
    cfg = dataclasses.replace(cases.pid_diag_cfg("float64", device="cuda"), use_cuda_graphs=True)
    a = S.FlowSolver(cfg)
    a.set_velocity(fx["u0"], fx["v0"])
    a.set_temperature(fx["theta0"])
    a.advance(5)
    sd = a.state_dict()
    a.set_bodies(new)
    assert a.bodies[0].name == "moved" and a.geometry.window != S.FlowSolver(cfg).geometry.window
    a.advance(15)
    b = S.FlowSolver(dataclasses.replace(cfg, bodies=new))
    b.load_state_dict(sd)
    b.advance(15)
    for f in ("u", "v", "p", "theta", "t", "dt", "ctrl"):
        assert torch.equal(getattr(a.state, f), getattr(b.state, f)), f
    assert np.array_equal(a.coefficients(), b.coefficients())
    ha = a.history()
    assert len(ha["t"]) == 20 and np.array_equal(ha["cd"][-15:], b.history()["cd"][-15:])
    assert np.array_equal(a.physics_history()["heat:moved"][-15:], b.physics_history()["heat:moved"][-15:])


def test_live_physics_edits_equal_fresh_solver():
    """set_emitters / set_thermal / set_sources on a running solver: the next steps equal a
    solver built with the edited configuration from the same state (smoke kept, theta kept)."""
    print("[pid gpu] live physics edits", flush=True)
    from aerolume_amd.ns2d.config import EmitterSpec, HeatSourceSpec
    fx = _fixture()
    e1 = EmitterSpec(position=(0.8, 1.0), kind="line", end=(0.8, 3.0), count=6, rate=1.0)
    e2 = EmitterSpec(position=(1.1, 1.4), kind="line", end=(1.1, 2.6), count=4, rate=0.8, pulse_period=0.2)
    heater = HeatSourceSpec(position=(3.2, 2.0), size=(0.3, 0.2), power=2.0)
    cfg = dataclasses.replace(cases.pid_diag_cfg("float64", device="cuda"), use_cuda_graphs=True, emitters=[e1])
    a = S.FlowSolver(cfg)
    a.set_velocity(fx["u0"], fx["v0"])
    a.set_temperature(fx["theta0"])
    a.advance(6)
    sd = a.state_dict()
    assert float(sd["dye"].abs().max()) > 0
    a.set_emitters([e2])
    a.set_thermal(richardson=0.3, theta_inlet=0.1)
    a.set_sources(heat_sources=[heater])
    a.advance(8)
    th = dataclasses.replace(cfg.thermal, richardson=0.3, theta_inlet=0.1)
    b = S.FlowSolver(dataclasses.replace(cfg, emitters=[e2], thermal=th, heat_sources=[heater]))
    b.load_state_dict(sd)
    b.advance(8)
    for f in ("u", "v", "p", "theta", "dye", "t", "dt", "ctrl"):
        assert torch.equal(getattr(a.state, f), getattr(b.state, f)), f
    assert a.physics_history().keys() == b.physics_history().keys()
    with pytest.raises(NotImplementedError):
        a.set_sources(fans=[object()])


@pytest.mark.parametrize("kind", ["periodic", "closed"])
def test_functional_diag_row_periodic_and_closed(kind):
    """The row's Poisson residual and divergence use the solve's own boundary rules: periodic
    wrap and closed walls (no outlet correction), against the oracle's row (float64)."""
    print(f"[pid gpu] diag row {kind}", flush=True)
    over = dict(thermal=None, bodies=[], adaptive_dt="cfl", turbulence="none", perturb=True)
    over.update(periodic=True) if kind == "periodic" else over.update(domain="closed")
    rows, states = [], []
    u0 = v0 = None
    for dev in ("cpu", "cuda"):
        cfg = cases.pid_diag_cfg("float64", device=dev, **over)
        ops = su.make_operators(cfg, device=dev)
        geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
        if u0 is None:
            u0, v0 = O.noisy_start(ops, 31, amplitude=0.5)
        st = S.init_state(cfg, ops)._replace(u=u0.to(ops.device), v=v0.to(ops.device))
        step = S.step_full if dev == "cuda" else O.step_full
        with torch.no_grad():
            st, _ = step(st, geom, ops)
            out = {}
            st, _ = step(st, geom, ops, out=out, diag=True)
        rows.append(out["diag"].double().cpu().numpy())
        states.append(st)
    c, g = rows
    assert float(states[0].u.abs().max()) > 0.1
    for i in range(18):
        if i in (3, 4, 5):
            assert g[i] < 1e-12 and c[i] < 1e-12, (kind, i, g[i], c[i])
        else:
            assert g[i] == pytest.approx(c[i], rel=1e-9, abs=1e-14), (kind, i, g[i], c[i])
    assert np.all(np.isnan(g[18:]))
