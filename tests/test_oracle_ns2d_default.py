"""CPU checks of the 2D oracle on the reference's default path (the Lamb-Oseen start and the
inlet smoke comb) and on the 2D variants no other fixture steps (central advection, Vreman,
WALE, no-slip and moving walls, parabolic inlet, doubly periodic box, closed lid-driven box),
against fixtures written from the reference's own FlowSolver and step_full
(``tests/golden/make_ns2d_default_fixtures.py``), plus the smoke scheme's invariants."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ns2d as O
from tests import cases


def _setup(cfg, live=False):
    from aerolume_amd.ns2d import setup as su
    ops = su.make_operators(cfg, device="cpu")
    w = su.live_window(cfg.bodies, ops) if live else None
    return ops, su.build_geometry(cfg.bodies, ops, window=w)


def _err(a, b):
    return float(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)).max())


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_oracle_init_state_matches_reference(golden, dtype):
    """The Lamb-Oseen start (core.py:216-271) equals the reference FlowSolver's initial state;
    the comb's smoke starts empty."""
    F = golden(f"ns2d_default_{'f64' if dtype == 'float64' else 'f32'}.npz")
    cfg = cases.default_cfg(dtype)
    ops, _ = _setup(cfg)
    st = O.init_state(cfg, ops)
    tol = 1e-14 if dtype == "float64" else 2e-7
    assert _err(st.u.numpy(), F["u_0"]) < tol and _err(st.v.numpy(), F["v_0"]) < tol
    assert float(np.abs(F["v_0"]).max()) > 1e-3          # the kick is really there
    assert st.dye is not None and float(st.dye.abs().max()) == 0.0 == float(np.abs(F["dye_0"]).max())
    # without the kick: the inlet profile, divergence-free as it stands
    plain = O.init_state(cfg, ops, perturb=False)
    assert float(plain.v.abs().max()) == 0.0 and float((plain.u - 1.0).abs().max()) == 0.0


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_oracle_default_path_matches_reference(golden, dtype):
    """20 steps of the default path from the reference's initial state: u, v, p, the inlet
    comb's dye and the coefficient history (FlowSolver uses the live body window)."""
    f64 = dtype == "float64"
    F = golden(f"ns2d_default_{'f64' if f64 else 'f32'}.npz")
    cfg = cases.default_cfg(dtype)
    ops, geom = _setup(cfg, live=True)
    to = lambda k: torch.as_tensor(F[k], dtype=ops.dtype)  # noqa: E731
    z = torch.zeros((), dtype=ops.dtype)
    cur = O.State(to("u_0"), to("v_0"), to("p_0"), to("dye_0"), z.clone(), z.clone())
    coeffs = []
    for n in range(1, 21):
        cur, c = O.step_full(cur, geom, ops)
        coeffs.append(c.numpy()[0])
        if n in (1, 10, 20):
            for k in ("u", "v", "dye"):
                assert _err(getattr(cur, k).numpy(), F[f"{k}_{n}"]) < (1e-13 if f64 else 3e-6), (k, n)
            assert _err(cur.p.numpy(), F[f"p_{n}"]) < (1e-11 if f64 else 5e-5), n
            assert _err(cur.t.numpy(), F[f"t_{n}"]) < (1e-15 if f64 else 1e-6)
    assert float(np.abs(F["dye_20"]).max()) > 0.5        # the comb's streaks have entered
    hist = np.stack([F["hist:cd"], F["hist:cl"], F["hist:cm"]], axis=1)
    assert _err(np.stack(coeffs), hist) < (1e-11 if f64 else 5e-5)


def test_inlet_comb_planted_mistake_is_caught(golden):
    """A comb that feeds zero smoke instead of its inlet column misses the reference's dye
    after one step by far more than round-off."""
    F = golden("ns2d_default_f64.npz")
    cfg = cases.default_cfg("float64")
    ops, geom = _setup(cfg, live=True)
    z = torch.zeros((), dtype=ops.dtype)
    st = O.State(torch.as_tensor(F["u_0"]), torch.as_tensor(F["v_0"]), torch.as_tensor(F["p_0"]),
                 torch.as_tensor(F["dye_0"]), z.clone(), z.clone())
    good, _ = O.step_full(st, geom, ops)
    assert _err(good.dye.numpy(), F["dye_1"]) < 1e-14
    saved = ops.dye_inlet
    try:
        ops.dye_inlet = torch.zeros_like(saved)
        bad, _ = O.step_full(st, geom, ops)
    finally:
        ops.dye_inlet = saved
    assert _err(bad.dye.numpy(), F["dye_1"]) > 1e-3


def test_advect_dye_invariants():
    """The comb's MUSCL step in a doubly periodic box: the total is conserved to round-off,
    no new extrema appear, and a uniform field stays uniform (core.py:709-733)."""
    from aerolume_amd.ns2d.config import SolverConfig
    cfg = SolverConfig(nx=32, ny=24, h=0.25, x0=0, y0=0, re=100, periodic=True, dtype="float64", device="cpu",
                       perturb=False)
    ops, _ = _setup(cfg)
    u, v = O.noisy_start(ops, 5, amplitude=0.3)
    rng = np.random.default_rng(9)
    c = torch.as_tensor(0.2 + 0.6 * rng.random((24, 32)))
    dt = 0.2 * ops.h / float(max(u.abs().max(), v.abs().max()))
    c1 = O.advect_dye(c, u, v, dt, ops)
    assert abs(float(c1.sum() - c.sum())) < 1e-12 * float(c.sum())
    assert float(c1.min()) >= float(c.min()) - 1e-14 and float(c1.max()) <= float(c.max()) + 1e-14
    flat = torch.full((24, 32), 0.375, dtype=torch.float64)
    assert _err(O.advect_dye(flat, u, v, dt, ops).numpy(), flat.numpy()) < 1e-15


@pytest.mark.parametrize("name", sorted(cases.VARIANTS2D))
def test_oracle_variants_match_reference(golden, name):
    """Five float64 steps of each 2D variant from a seeded noisy start equal the reference's."""
    F = golden("ns2d_variants.npz")
    cfg = cases.variant_cfg(name)
    ops, geom = _setup(cfg)
    z = torch.zeros((), dtype=ops.dtype)
    cur = O.State(torch.as_tensor(F[f"{name}:u0"]), torch.as_tensor(F[f"{name}:v0"]),
                  torch.zeros(ops.ny, ops.nx, dtype=ops.dtype), None, z.clone(), z.clone())
    cs = []
    for _ in range(5):
        cur, c = O.step_full(cur, geom, ops)
        cs.append(c.numpy())
    scale = float(np.abs(F[f"{name}:u5"]).max())
    assert _err(cur.u.numpy(), F[f"{name}:u5"]) < 1e-12 * scale
    assert _err(cur.v.numpy(), F[f"{name}:v5"]) < 1e-12 * scale
    assert _err(cur.p.numpy(), F[f"{name}:p5"]) < 1e-10 * max(1.0, float(np.abs(F[f"{name}:p5"]).max()))
    assert _err(cur.t.numpy(), F[f"{name}:t5"]) < 1e-14
    if cs[0].size:
        assert _err(np.stack(cs), F[f"{name}:coeffs"]) < 1e-10


def test_wall_shear_excess_solves_the_power_law():
    """The oracle's Werner-Wengle stress against the power law solved numerically: for a face
    velocity us at distance y, find u_tau with us / u_tau = A (u_tau y / nu)^B (a root finder,
    not the closed form), then the excess is max(u_tau^2 - nu |us| / y, 0) with the sign of us;
    zero in the viscous sublayer (us y / nu below A^(2/(1-B)))."""
    from scipy.optimize import brentq
    nu, y = 1e-4, 0.03
    for us in (-2.0, -0.3, -1e-3, 0.0, 2e-4, 5e-3, 0.07, 1.3, 9.0):
        got = float(O.wall_shear_excess(torch.tensor(us, dtype=torch.float64), y, torch.tensor(nu,
                                                                                                dtype=torch.float64)))
        a = abs(us)
        if a == 0.0:
            assert got == 0.0
            continue
        ut = brentq(lambda t: a / t - O.WW_A * (t * y / nu) ** O.WW_B, 1e-12, 1e3, xtol=1e-300, rtol=1e-15)
        want = np.sign(us) * max(ut * ut - nu * a / y, 0.0)
        assert abs(got - want) <= 1e-12 * max(abs(want), nu * a / y), us
        sublayer = a * y / nu < O.WW_A ** (2.0 / (1.0 - O.WW_B))
        assert (got == 0.0) == sublayer, us


@pytest.mark.parametrize("name", ["ww_tunnel_moving", "ww_closed_room"])
def test_wall_model_acts_in_its_fixture_cases(golden, name):
    """The wall-model variants differ from the same runs without the model (so the fixture
    comparison pins the model, not just the rest of the step), and the difference starts on
    the modelled face rows: the tendencies of the reference's state after five steps change on
    the wall rows only (and on both walls)."""
    import dataclasses
    F = golden("ns2d_variants.npz")
    cfg = cases.variant_cfg(name)
    ops, geom = _setup(cfg)
    off = dataclasses.replace(ops, wall_model=False)
    u5, v5 = torch.as_tensor(F[f"{name}:u5"]), torch.as_tensor(F[f"{name}:v5"])
    lu1, lv1, _, _ = O.momentum_rhs(u5, v5, geom, ops)
    lu0, lv0, _, _ = O.momentum_rhs(u5, v5, geom, off)
    rows = (lu1 - lu0).abs().amax(dim=1)
    assert float(rows[0]) > 0 and float(rows[-1]) > 0 and float(rows[1:-1].abs().max()) == 0.0
    cols = (lv1 - lv0).abs().amax(dim=0)
    if ops.closed:
        assert float(cols[0]) > 0 and float(cols[-1]) > 0 and float(cols[1:-1].max()) == 0.0
    else:
        assert float(cols.max()) == 0.0
