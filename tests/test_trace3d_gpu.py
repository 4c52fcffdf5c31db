"""GPU parity of the traced 3D step (``FlowSolver3D.trace`` / ``trace_step`` through
``aero3d_step_trace``) against the reference's own stage records
(``tests/golden/ns3d_trace_*.npz``, float64): the staggered stage arrays to round-off
(1e-10 relative), the cell-centred float32 outputs of ``trace()`` to float32 rounding, the
flat keys of ``trace_step``, and that tracing leaves the run untouched."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
TAGS = ("tend", "star", "pen", "proj")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _fixture(name):
    path = GOLDEN / name
    if not path.exists():
        pytest.skip(f"missing fixture {name}")
    return dict(np.load(path))


def _solver(case):
    from aerolume_amd.ns3d import solver as S
    if case == "wale_moving_py":
        start = _fixture("ns3d_wale_moving_py_f64.npz")
        s = S.FlowSolver3D(cases.cfg3d(case, "float64", device="cuda"))
    else:
        start = _fixture("ns3d_thermal_box.npz")
        s = S.FlowSolver3D(cases.thermal_cfg("box", start, "float64", device="cuda"))
        s.set_temperature(start["theta0"])
    s.set_velocity(start["u0"], start["v0"], start["w0"])
    return s


def _rel(a, b):
    a = a.detach().double().cpu().numpy() if torch.is_tensor(a) else np.asarray(a, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1.0))


def _centre(x, axis):
    return 0.5 * (np.take(x, range(1, x.shape[axis]), axis=axis) + np.take(x, range(x.shape[axis] - 1), axis=axis))


@pytest.mark.parametrize("case", ["wale_moving_py", "box"])
def test_traced_stages_match_reference_records(case):
    _need_gpu()
    print(f"[trace3d gpu] staggered records {case}", flush=True)
    ref = _fixture(f"ns3d_trace_{case}.npz")
    s = _solver(case)
    run = s._traced_step()
    torch.cuda.synchronize()
    st = run["stages"]
    assert float(run["new"]["dt"]) == pytest.approx(float(ref["dt"]), rel=1e-13)
    for k in range(3):
        for tag in TAGS:
            for c in "uvw":
                assert _rel(st[f"{tag}_{c}"][k], ref[f"s{k}_{tag}_{c}"]) < 1e-10, (k, tag, c)
        assert _rel(st["div"][k], ref[f"s{k}_rhs"]) < 1e-10, k
        assert _rel(st["psi"][k], ref[f"s{k}_psi"]) < 1e-10, k
        if case == "box":
            assert _rel(st["theta"][k], ref[f"s{k}_theta"]) < 1e-10, k


@pytest.mark.parametrize("case", ["wale_moving_py", "box"])
def test_trace_and_trace_step_outputs(case):
    _need_gpu()
    print(f"[trace3d gpu] trace / trace_step {case}", flush=True)
    ref = _fixture(f"ns3d_trace_{case}.npz")
    s = _solver(case)
    before = {c: getattr(s.state, c).clone() for c in "uvw"}
    t0 = s.t
    tr = s.trace()
    assert tr["dt"] == pytest.approx(float(ref["dt"]), rel=1e-13)
    assert [st["stage"] for st in tr["stages"]] == [0, 1, 2]
    for k, st in enumerate(tr["stages"]):
        for key, tag in (("tendency", "tend"), ("u_star", "star"), ("u_penalised", "pen"), ("u_projected", "proj")):
            for axis, c, x in zip((2, 1, 0), "uvw", st[key]):
                assert x.dtype == torch.float32 and x.device.type == "cpu"
                want = _centre(ref[f"s{k}_{tag}_{c}"], axis)
                assert _rel(x, want) < 2e-7, (k, key, c)
        assert _rel(st["divergence"], ref[f"s{k}_rhs"]) < 2e-7
        assert _rel(st["psi"], ref[f"s{k}_psi"]) < 2e-7
        assert st["max_rhs"] == pytest.approx(float(np.abs(ref[f"s{k}_rhs"]).max()), rel=1e-10)
        assert st["max_div_after"] < 1e-10 * max(1.0, st["max_rhs"])
        assert st["poisson_residual"] < 1e-10
        assert st["correction_max"] > 0.0
    # tracing leaves the run as it was
    for c in "uvw":
        assert torch.equal(getattr(s.state, c), before[c])
    assert s.t == t0 and s.steps == 0
    flat = s.trace_step(commit=True)
    names = {f"stage{k}/{n}_{c}" for k in (1, 2, 3) for n in ("explicit", "penalised", "correction", "projected")
             for c in "uvw"} | {f"stage{k}/{n}" for k in (1, 2, 3) for n in ("divergence", "pressure_potential")}
    names |= {"final_u", "final_v", "final_w", "final_p"}
    if case == "box":
        names |= {f"stage{k}/theta" for k in (1, 2, 3)} | {"final_theta"}
    assert set(flat) == names
    # the committed step is the traced one
    assert s.steps == 1
    for axis, c in zip((2, 1, 0), "uvw"):
        got = getattr(s.state, c).double().cpu().numpy()
        assert _rel(flat[f"final_{c}"], _centre(got, axis)) < 2e-7
    p = ref["p"]
    shift = p.mean() if (s.ops.periodic_axes[0] or s.ops.closed) else p[..., :2].mean()
    assert _rel(flat["final_p"], p - shift) < 2e-6


def test_trace_then_advance_equals_plain_advance():
    """A trace between steps changes nothing: the next steps equal those of an untraced run
    bit for bit (graphs on)."""
    _need_gpu()
    print("[trace3d gpu] trace between steps", flush=True)
    a, b = _solver("wale_moving_py"), _solver("wale_moving_py")
    a.advance(12)
    b.advance(12)
    a.trace()
    a.advance(12)
    b.advance(12)
    for c in "uvwp":
        assert torch.equal(getattr(a.state, c), getattr(b.state, c)), c
    assert np.array_equal(a.history()["cd"], b.history()["cd"])
