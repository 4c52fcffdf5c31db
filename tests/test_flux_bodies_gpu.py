"""GPU parity of 2D heat-flux bodies (skin sources and skin temperatures through the C ABI)
against the reference's own run (``tests/golden/ns2d_flux_bodies.npz``, float64, to round-off)
and the oracle (float32, within the tolerances written below), plus the solver-level reports:
the aux row layout, ``heat_transfer`` (G P / Pe for the flux body) and graph replays."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _err(a, b):
    return float(np.abs(np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
                        - np.asarray(b, dtype=np.float64)).max())


def _setup(dtype, device):
    from aerolume_amd.ns2d import setup as su
    cfg = cases.flux_cfg(dtype, device="cuda" if device.type == "cuda" else "cpu")
    ops = su.make_operators(cfg, device=device)
    geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
    return cfg, ops, geom


def test_flux_bodies_match_reference_float64(golden):
    """20 float64 steps of the functional step: fields to 1e-11, the held disc's heat and the
    flux disc's skin temperature to 1e-10 relative, walls, storage and flow to round-off."""
    _need_gpu()
    from aerolume_amd.ns2d import solver as S
    print("[flux gpu] float64 vs reference", flush=True)
    F = golden("ns2d_flux_bodies.npz")
    cfg, ops, geom = _setup("float64", DEV)
    assert tuple(geom.window) == tuple(int(x) for x in F["window"])
    t = lambda a: torch.as_tensor(a, device=DEV, dtype=ops.dtype)  # noqa: E731
    z = torch.zeros((), device=DEV, dtype=ops.dtype)
    cur = S.State(t(F["u0"]), t(F["v0"]), torch.zeros(ops.ny, ops.nx, device=DEV, dtype=ops.dtype), None,
                  z.clone(), z.clone(), t(F["theta0"]))
    for n in range(1, 21):
        out = {}
        cur, c = S.step_full(cur, geom, ops, out=out)
        assert abs(float(cur.dt) - F["hist_dt"][n - 1]) < 1e-15
        assert _err(c, F["hist_coeffs"][n - 1]) < 1e-9
        for k in ("heat", "theta_wall", "wall_heat", "heat_storage", "flow_heat"):
            want = F["hist_" + k][n - 1]
            assert _err(out[k], want) <= 1e-10 * max(1.0, np.abs(want).max()), (n, k)
        if n in (1, 20):
            for k in ("u", "v", "theta"):
                assert _err(getattr(cur, k), F[f"{k}{n}"]) < 1e-11, (k, n)
            assert _err(cur.p, F[f"p{n}"]) < 1e-9, n


def test_flux_bodies_float32_match_oracle(golden):
    """float32 on the GPU against the float32 oracle from the same start, 20 steps: theta and
    the skin temperature to 2e-6, u, v to 5e-5."""
    _need_gpu()
    import oracle.ns2d as O
    from aerolume_amd.ns2d import solver as S
    print("[flux gpu] float32 vs oracle", flush=True)
    F = golden("ns2d_flux_bodies.npz")
    _, ops, geom = _setup("float32", DEV)
    _, ops_c, geom_c = _setup("float32", torch.device("cpu"))
    t = lambda a, d: torch.as_tensor(a, device=d, dtype=torch.float32)  # noqa: E731
    z = lambda d: torch.zeros((), device=d, dtype=torch.float32)  # noqa: E731
    cur = S.State(t(F["u0"], DEV), t(F["v0"], DEV), torch.zeros(ops.ny, ops.nx, device=DEV), None, z(DEV), z(DEV),
                  t(F["theta0"], DEV))
    ref = O.State(t(F["u0"], "cpu"), t(F["v0"], "cpu"), torch.zeros(ops.ny, ops.nx), None, z("cpu"), z("cpu"),
                  t(F["theta0"], "cpu"))
    for _ in range(20):
        out, out_c = {}, {}
        cur, _ = S.step_full(cur, geom, ops, out=out)
        ref, _ = O.step_full(ref, geom_c, ops_c, out=out_c)
    assert _err(out["theta_wall"], out_c["theta_wall"].numpy()) < 2e-6
    assert _err(cur.theta, ref.theta.numpy()) < 2e-6
    assert _err(cur.u, ref.u.numpy()) < 5e-5 and _err(cur.v, ref.v.numpy()) < 5e-5


def test_flowsolver_reports_flux_bodies(golden):
    """FlowSolver with a heat-flux body: the aux columns put the skin temperature after the
    held body's heat (the reference's layout), graphs replay the eager steps bit for bit, and
    heat_transfer gives the flux body G P / Pe with its recorded skin temperature."""
    _need_gpu()
    import dataclasses
    from aerolume_amd.ns2d import solver as S
    print("[flux gpu] FlowSolver reports", flush=True)
    F = golden("ns2d_flux_bodies.npz")
    cfg = cases.flux_cfg("float64", device="cuda")
    runs = []
    for graphs in (True, False):
        s = S.FlowSolver(dataclasses.replace(cfg, use_cuda_graphs=graphs))
        s.set_velocity(F["u0"], F["v0"])
        s.set_temperature(F["theta0"])
        s.advance(20)
        runs.append(s)
    a, b = runs
    assert torch.equal(a.state.theta, b.state.theta) and torch.equal(a.state.u, b.state.u)
    ph = a.physics_history()
    assert list(ph)[:4] == ["t", "heat:held", "theta_wall:flux", "heat:left"]
    assert np.array_equal(ph["theta_wall:flux"], b.physics_history()["theta_wall:flux"])
    assert np.abs(ph["theta_wall:flux"] - F["hist_theta_wall"][:, 0]).max() < 1e-10
    assert np.abs(ph["heat:held"] - F["hist_heat"][:, 0]).max() < 1e-9 * np.abs(F["hist_heat"]).max()
    rep = a.heat_transfer()
    flux = [r for r in rep["bodies"] if r["name"] == "flux"][0]
    pe = cfg.re * cfg.thermal.prandtl
    assert flux["heat_rate"] == pytest.approx(2.0 * float(F["g_perimeter"][0]) / pe, rel=1e-12)
    assert flux["theta_wall"] == pytest.approx(float(F["hist_theta_wall"][-1, 0]), rel=1e-10)
    bal = a.heat_balance()
    assert bal["heat_in_bodies"] == pytest.approx(flux["heat_rate"] + float(ph["heat:held"][-1]), rel=1e-12)
