"""The 2D PID-controlled step (``adaptive_dt="pid"``) on the MI355X through the C ABI: the
reference's trajectory in float64 (``tests/golden/ns2d_pid_diag.npz``), the oracle in float32,
and the FlowSolver's graphs and chained steps against eager plain steps."""
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
