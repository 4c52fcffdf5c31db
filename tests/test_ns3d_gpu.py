"""GPU parity of the 3D step: HIP kernels (through the C ABI) against the CPU oracle and the
reference's golden trajectories. Bars: float64 to round-off (1e-10 relative or tighter),
float32 within the tolerances written in each test."""
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


def _S():
    from aerolume_amd.ns3d import solver as S
    return S


def _err(a, b):
    return float(np.abs(np.asarray(a.detach().cpu() if torch.is_tensor(a) else a, dtype=np.float64)
                        - np.asarray(b, dtype=np.float64)).max())


def _ops(cfg):
    return _S().make_operators3d(cfg, device=DEV)


def _cpu_ops(cfg):
    from aerolume_amd.ns3d import setup as su
    return su.make_operators3d(cfg, device="cpu")


def _state(ops, u, v, w):
    S = _S()
    z = torch.zeros((), device=DEV, dtype=ops.dtype)
    to = lambda a: torch.as_tensor(np.asarray(a), device=DEV, dtype=ops.dtype).contiguous()  # noqa: E731
    return S.State3D(to(u), to(v), to(w), torch.zeros(ops.nz, ops.ny, ops.nx, device=DEV, dtype=ops.dtype),
                     z.clone(), z.clone())


def _cfg(nx, ny, nz, dtype, **kw):
    from aerolume_amd.ns3d.config import SolverConfig3D
    return SolverConfig3D(nx=nx, ny=ny, nz=nz, h=0.1, x0=0, y0=0, z0=0, re=100.0, dtype=dtype, device="cuda",
                          dye=False, perturb=False, **kw)


@pytest.mark.parametrize("shape,kw,dtype", [((24, 16, 16), {}, "float64"), ((30, 20, 12), {"bc_y": "periodic"}, "float64"),
                                            ((16, 16, 16), {"periodic": True}, "float64"),
                                            ((40, 24, 20), {"bc_y": "periodic"}, "float32"),
                                            ((32, 16, 24), {}, "float32"),
                                            ((16, 16, 16), {"periodic": True}, "float32")])
def test_poisson_solve_matches_oracle(shape, kw, dtype):
    """Exact DCT / DHT solve on the GPU equals the oracle's dense-basis solve and satisfies
    L psi = f - mean(f). The float32 cases take the fused (z, y) plane pass (even ny and nz),
    float64 the separate y and z passes."""
    _need_gpu()
    from oracle import ns3d as O
    S = _S()
    nx, ny, nz = shape
    print(f"poisson3d {shape} {kw} {dtype}", flush=True)
    cfg = _cfg(nx, ny, nz, dtype, **kw)
    ops = _ops(cfg)
    f = torch.as_tensor(np.random.default_rng(3).standard_normal((nz, ny, nx)), dtype=ops.dtype)
    psi = S.poisson_solve(f.to(DEV), ops).cpu()
    ref = O.poisson(_cpu_ops(_cfg(nx, ny, nz, dtype, **kw))).solve(f)
    tol = 1e-11 if dtype == "float64" else 2e-5
    scale = float(ref.abs().max())
    assert _err(psi, ref) < tol * scale
    ps = O.PoissonSolver3D(nx, ny, nz, 0.1, periodic=ops.periodic_axes, dtype=torch.float64)
    r = ps.laplacian(psi.double()) - (f.double() - f.double().mean())
    assert float(r.abs().max()) < (1e-9 if dtype == "float64" else 2e-3) * float(f.abs().max())


@pytest.mark.parametrize("name", ["laminar", "wale_moving_py", "closed_slip", "periodic_wale"])
def test_projection_matches_oracle(name):
    _need_gpu()
    from oracle import ns3d as O
    S = _S()
    print(f"project3d {name}", flush=True)
    cfg = cases.cfg3d(name, "float64", device="cuda")
    ops = _ops(cfg)
    cops = _cpu_ops(cases.cfg3d(name, "float64"))
    rng = np.random.default_rng(5)
    u = torch.as_tensor(rng.standard_normal((ops.nz, ops.ny, ops.nx + 1)))
    v = torch.as_tensor(rng.standard_normal((ops.nz, ops.ny + 1, ops.nx)))
    w = torch.as_tensor(rng.standard_normal((ops.nz + 1, ops.ny, ops.nx)))
    u, v, w = O.apply_boundary_values(u, v, w, cops)
    ru, rv, rw, rpsi = O.project3d(u, v, w, cops)
    gu, gv, gw, gpsi = S.project3d(u.to(DEV).contiguous(), v.to(DEV).contiguous(), w.to(DEV).contiguous(), ops)
    for a, b in ((gu, ru), (gv, rv), (gw, rw), (gpsi, rpsi)):
        assert _err(a, b) < 1e-11 * max(1.0, float(b.abs().max()))
    assert float(S.divergence3d(gu, gv, gw, ops.h).abs().max()) * ops.h < 1e-10


@pytest.mark.parametrize("name", list(cases.CASES3D))
def test_steps_match_reference_float64(name, golden):
    """Five float64 HIP steps from the reference's seeded start against the reference's own
    trajectory: fields, dt, t and the six coefficients per body to round-off."""
    _need_gpu()
    S = _S()
    from aerolume_amd.ns3d import setup as su
    print(f"step3d f64 {name}", flush=True)
    fx = golden(f"ns3d_{name}_f64.npz")
    cfg = cases.cfg3d(name, "float64", device="cuda")
    ops = _ops(cfg)
    g = su.build_geometry3d(cfg.bodies, ops)
    st = _state(ops, fx["u0"], fx["v0"], fx["w0"])
    for n in range(1, 6):
        st, c = S.step3d_full(st, g, ops)
        assert abs(float(st.dt) - fx["dt"][n - 1]) <= 1e-12 * fx["dt"][n - 1], n
        if c.numel():
            assert _err(c, fx["coeffs"][n - 1]) < 1e-9 * max(1.0, np.abs(fx["coeffs"][n - 1]).max()), n
        if n in (1, 5):
            for f in "uvwp":
                ref = fx[f"{f}_{n}"]
                assert _err(getattr(st, f), ref) < 1e-10 * max(1.0, np.abs(ref).max()), (f, n)
    assert abs(float(st.t) - fx["t"][-1]) < 1e-12


def test_car_miniature_float32_matches_reference(golden):
    """The car miniature (WALE, moving floor, periodic y, five bodies, wheels rolling) in
    float32 for 20 steps against the reference's float32 run: |du| < 2e-5, |dp| < 1e-4,
    coefficients within 1e-3."""
    _need_gpu()
    S = _S()
    from aerolume_amd.ns3d import setup as su
    fx = golden("ns3d_wale_moving_py_f32.npz")
    cfg = cases.cfg3d("wale_moving_py", "float32", device="cuda")
    ops = _ops(cfg)
    g = su.build_geometry3d(cfg.bodies, ops)
    st = _state(ops, fx["u0"], fx["v0"], fx["w0"])
    for _ in range(20):
        st, c = S.step3d_full(st, g, ops)
    for f in "uvw":
        assert _err(getattr(st, f), fx[f"{f}_20"]) < 2e-5, f
    assert _err(st.p, fx["p_20"]) < 1e-4
    assert _err(c, fx["coeffs"][-1]) < 1e-3


def test_explicit_dt_and_fixed_dt():
    """A dt passed to the step is used as given (and leaves the CFL plan alone)."""
    _need_gpu()
    from oracle import ns3d as O
    from aerolume_amd.ns3d import setup as su
    S = _S()
    cfg = cases.cfg3d("laminar", "float64", device="cuda")
    ops = _ops(cfg)
    g = su.build_geometry3d(cfg.bodies, ops)
    cops = _cpu_ops(cases.cfg3d("laminar", "float64"))
    cg = su.build_geometry3d(cfg.bodies, cops)
    u, v, w = O.noisy_start3d(cops, 8)
    st, c = S.step3d_full(_state(ops, u, v, w), g, ops, dt=0.01)
    z = torch.zeros((), dtype=torch.float64)
    ref, rc = O.step3d_full(O.State3D(u, v, w, torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z, z),
                            cg, cops, dt=0.01)
    assert abs(float(st.dt) - 0.01) < 1e-15
    for f in "uvwp":
        assert _err(getattr(st, f), getattr(ref, f)) < 1e-10 * max(1.0, float(getattr(ref, f).abs().max()))
    assert _err(c, rc) < 1e-9


def test_flow_solver_graphs_equal_eager_and_chained_equal_plain():
    """FlowSolver3D: HIP-graph replays give bit-identical states to eager steps, and the
    chained hot loop equals plain steps that re-read their dt inputs."""
    _need_gpu()
    S = _S()
    import dataclasses
    base = cases.cfg3d("wale_moving_py", "float32", device="cuda")
    runs = {}
    for mode in ("graphs", "eager", "plain"):
        cfg = dataclasses.replace(base, use_cuda_graphs=(mode == "graphs"), perturb=True)
        fs = S.FlowSolver3D(cfg)
        if mode == "plain":
            for _ in range(25):
                fs._touched()
                fs.advance(1)
        else:
            fs.advance(25)
        torch.cuda.synchronize()
        runs[mode] = fs
    a, b, c = runs["graphs"].state, runs["eager"].state, runs["plain"].state
    for f in "uvwp":
        assert torch.equal(getattr(a, f), getattr(b, f)), f
        assert torch.equal(getattr(b, f), getattr(c, f)), f
    h = runs["graphs"].history()
    assert len(h["t"]) == 25 and np.all(np.diff(h["t"]) > 0)
    assert np.allclose(h["cd"][-1], runs["graphs"].coefficients()[0])
    for name in _S().FIELDS3D:
        fld = runs["graphs"].field(name)
        assert fld.shape == (base.nz, base.ny, base.nx) and bool(torch.isfinite(fld).all()), name


def test_full_size_car_one_step_matches_oracle():
    """The 256x128x128 car case (4.2M cells, FP32): one HIP step from a seeded noisy start
    against the CPU oracle's step: |du| < 5e-5, |dp| < 5e-4, coefficients within 1e-3, and
    the step leaves a discretely divergence-free field (|div| h < 1e-5)."""
    _need_gpu()
    from oracle import ns3d as O
    from aerolume_amd.ns3d import setup as su
    S = _S()
    print("full-size car step", flush=True)
    cfg = cases.c3_cfg("float32", device="cuda")
    ops = _ops(cfg)
    cops = _cpu_ops(cases.c3_cfg("float32", device="cpu"))
    cg = su.build_geometry3d(cfg.bodies, cops)
    g = cg.to(DEV, ops.dtype)
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    u, v, w = O.noisy_start3d(cops, 303)
    z = torch.zeros((), dtype=torch.float32)
    ref, rc = O.step3d_full(O.State3D(u, v, w, torch.zeros(ops.nz, ops.ny, ops.nx), z, z), cg, cops)
    st, c = S.step3d_full(_state(ops, u, v, w), g, ops)
    assert abs(float(st.dt) - float(ref.dt)) < 1e-6 * float(ref.dt)
    for f in "uvw":
        assert _err(getattr(st, f), getattr(ref, f)) < 5e-5, f
    assert _err(st.p, ref.p) < 5e-4
    assert _err(c, rc) < 1e-3
    assert float(S.divergence3d(st.u, st.v, st.w, ops.h).abs().max()) * ops.h < 1e-5


@pytest.mark.parametrize("name", ["laminar", "wale_moving_py"])
def test_flow_solver_smoke_matches_oracle(name):
    """FlowSolver3D with smoke (automatic rake; pulsed on the car case), FP64, against the
    oracle's flow steps followed by its smoke step (MUSCL advection with the end-of-step
    velocity, solids, sources at the new time), 6 steps."""
    _need_gpu()
    import dataclasses
    from oracle import ns3d as O
    from aerolume_amd.ns3d import setup as su
    from aerolume_amd.ns3d.config import DyeRake
    S = _S()
    cfg = cases.cfg3d(name, "float64", device="cuda")
    cops = _cpu_ops(cases.cfg3d(name, "float64"))
    rake = su.auto_dye_rake(cfg, cops)
    if name == "wale_moving_py":
        rake = DyeRake(rake.kind, rake.positions, rake.radius, pulse_on=0.05, pulse_off=0.03)
    cfg = dataclasses.replace(cfg, dye=True, dye_rake=rake, use_cuda_graphs=False)
    fs = S.FlowSolver3D(cfg)
    u, v, w = O.noisy_start3d(cops, 17)
    fs.set_velocity(u, v, w)
    cg = su.build_geometry3d(cfg.bodies, cops)
    win, wt = su.dye_source_weights(rake, cops)
    pulse = torch.tensor(su.dye_pulse(rake), dtype=torch.float64)
    z = torch.zeros((), dtype=torch.float64)
    st = O.State3D(u, v, w, torch.zeros(cops.nz, cops.ny, cops.nx, dtype=torch.float64), z, z)
    c = torch.zeros(cops.nz, cops.ny, cops.nx, dtype=torch.float64)
    for n in range(6):
        st, _ = O.step3d_full(st, cg, cops)
        c = O.advect_dye3d(c, st.u, st.v, st.w, st.dt, cops)
        O.apply_solids_and_sources(c, st.t, cg, win, torch.as_tensor(wt), pulse)
    fs.advance(6)
    assert _err(fs.state.u, st.u) < 1e-10
    assert float(c.max()) > 0.5
    assert _err(fs.field("dye"), c) < 1e-10


def test_set_dye_rake_equals_fresh_solver():
    """A rake moved live (and pulsed) gives the same smoke as a solver configured with that
    rake and started from the same state (FP64, graphs on)."""
    _need_gpu()
    import dataclasses
    from oracle import ns3d as O
    from aerolume_amd.ns3d.config import DyeRake
    S = _S()
    cfg = dataclasses.replace(cases.cfg3d("laminar", "float64", device="cuda"), dye=True, use_cuda_graphs=True)
    a = S.FlowSolver3D(cfg)
    assert a.dye_channels == 1
    auto = a.default_dye_rake()
    assert np.allclose(np.asarray(auto.positions), np.asarray(a.dye_rake.positions))
    cops = _cpu_ops(cases.cfg3d("laminar", "float64"))
    a.set_velocity(*O.noisy_start3d(cops, 17))
    a.advance(4)
    sd = a.state_dict()
    moved = DyeRake(auto.kind, np.asarray(auto.positions) + np.array([0.1, 0.0, 0.05]), auto.radius,
                    pulse_on=0.04, pulse_off=0.02)
    a.set_dye_rake(moved)
    a.advance(5)
    b = S.FlowSolver3D(dataclasses.replace(cfg, dye_rake=moved))
    b.load_state_dict(sd)
    b.advance(5)
    assert torch.equal(a.field("dye"), b.field("dye"))
    assert torch.equal(a.state.u, b.state.u)
