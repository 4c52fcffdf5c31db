"""CPU checks of the oracle's traced 3D step (the per-stage records of core.py:822-862) against
the reference's own records (``tests/golden/ns3d_trace_*.npz``), plus what the records must
satisfy by construction: the penalised value equals the value before penalisation outside the
body window, the projected velocity is the penalised one less grad psi and is divergence-free,
and the step's pressure is the weighted sum of the stage potentials over dt."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ns3d as O
from tests import cases

TAGS = ("tend", "star", "pen", "proj")
KEYS = ("tendency", "u_star", "u_penalised", "u_projected")


def _records_car(golden):
    from aerolume_amd.ns3d import setup as su
    fx = golden("ns3d_wale_moving_py_f64.npz")
    cfg = cases.cfg3d("wale_moving_py", "float64")
    ops = su.make_operators3d(cfg, device="cpu")
    g = su.build_geometry3d(cfg.bodies, ops)
    z = torch.zeros((), dtype=torch.float64)
    st = O.State3D(*(torch.as_tensor(fx[k]) for k in ("u0", "v0", "w0")),
                   torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z.clone(), z.clone())
    rec: list = []
    new, _ = O.step3d_full(st, g, ops, trace=rec)
    return rec, new, ops, g


def _records_box(golden):
    from aerolume_amd.ns3d import setup as su
    from aerolume_amd.ns3d import thermal as T
    fx = golden("ns3d_thermal_box.npz")
    cfg = cases.thermal_cfg("box", fx, "float64")
    ops = su.make_operators3d(cfg, device="cpu")
    g = su.build_geometry3d(cfg.bodies, ops)
    th = O.Thermal(T.build_thermal(cfg, ops, cfg.bodies, g.window), ops, cfg.re)
    z = torch.zeros((), dtype=torch.float64)
    st = O.State3D(*(torch.as_tensor(fx[k]) for k in ("u0", "v0", "w0")),
                   torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z.clone(), z.clone())
    rec: list = []
    new, _, _ = O.step3d_phys(st, g, ops, th, torch.as_tensor(fx["theta0"]), trace=rec)
    return rec, new, ops, g


@pytest.mark.parametrize("case", ["wale_moving_py", "box"])
def test_oracle_trace_matches_reference_records(case, golden):
    ref = golden(f"ns3d_trace_{case}.npz")
    rec, new, ops, g = (_records_car if case == "wale_moving_py" else _records_box)(golden)
    assert len(rec) == 3
    assert float(rec[0]["dt"]) == pytest.approx(float(ref["dt"]), rel=1e-14)
    for r in rec:
        s = r["stage"]
        for key, tag in zip(KEYS, TAGS):
            for c, x in zip("uvw", r[key]):
                want = ref[f"s{s}_{tag}_{c}"]
                assert np.abs(x.numpy() - want).max() <= 1e-11 * max(1.0, np.abs(want).max()), (s, tag, c)
        for k in ("rhs", "psi") + (("theta",) if "theta" in r else ()):
            want = ref[f"s{s}_{k}"]
            assert np.abs(r[k].numpy() - want).max() <= 1e-11 * max(1.0, np.abs(want).max()), (s, k)
    assert ("s0_theta" in ref) == ("theta" in rec[0])
    assert np.abs(new.p.numpy() - ref["p"]).max() <= 1e-10 * max(1.0, np.abs(ref["p"]).max())


@pytest.mark.parametrize("case", ["wale_moving_py", "box"])
def test_oracle_trace_records_are_consistent(case, golden):
    rec, new, ops, g = (_records_car if case == "wale_moving_py" else _records_box)(golden)
    h = ops.h
    for r in rec:
        # penalisation only acts on the body window: elsewhere the two values agree exactly
        # (the buoyancy impulse of the heated case moves w everywhere, so only u, v are checked there)
        k0, k1, j0, j1, i0, i1 = g.window
        comps = "uv" if case == "box" else "uvw"
        for c, star, pen in zip("uvw", r["u_star"], r["u_penalised"]):
            if c not in comps:
                continue
            mask = torch.ones_like(star, dtype=torch.bool)
            mask[k0:k1 + (c == "w"), j0:j1 + (c == "v"), i0:i1 + (c == "u")] = False
            assert torch.equal(star[mask], pen[mask]), c
        gx, gy, gz = O.pressure_gradient(r["psi"], ops)
        for pen, proj, grad in zip(r["u_penalised"], r["u_projected"], (gx, gy, gz)):
            assert torch.equal(pen - grad, proj)
        div = O.divergence3d(*r["u_projected"], h)
        assert float(div.abs().max()) < 1e-10 * max(1.0, float(r["rhs"].abs().max()))
        assert torch.equal(r["rhs"], O.divergence3d(*r["u_penalised"], h))
    p = sum(w * r["psi"] for w, r in zip(O.FORCE_WEIGHTS, rec)) * (1.0 / rec[0]["dt"])
    assert torch.allclose(p, new.p, rtol=0, atol=1e-12 * max(1.0, float(p.abs().max())))
    for c, proj in zip("uvw", rec[-1]["u_projected"]):
        assert torch.equal(proj, getattr(new, c))


def test_planted_trace_mistake_is_caught(golden):
    """A trace that recorded the tendency before the boundary rules (a plausible slip) fails
    the reference comparison."""
    ref = golden("ns3d_trace_wale_moving_py.npz")
    rec, _, _, _ = _records_car(golden)
    lu = rec[0]["tendency"][0].clone()
    lu[..., -1] = 0.0          # the convective outlet tendency dropped
    assert np.abs(lu.numpy() - ref["s0_tend_u"]).max() > 1e-6
