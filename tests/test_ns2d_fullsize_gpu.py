"""GPU parity of the reference's default 2D path, the 2D variants, the heated-airfoil case at
its own 1024x512 size and the 64-run batched sweep, through the C ABI.

* the default path (``FlowSolver(cfg)`` with every SolverConfig default: Lamb-Oseen start,
  inlet smoke comb) against the reference FlowSolver's fixture;
* five float64 steps of each 2D variant (central + Vreman + no-slip walls + parabolic inlet,
  WALE over a moving wall, doubly periodic Smagorinsky, closed lid-driven box) against the
  reference's step_full;
* the 1024x512 heated airfoil (1024- and 512-thread transform blocks, XCD-ordered column
  pass) against the CPU oracle at steps 1, 10 and 100, FP32 and FP64;
* the 64-run sweep on 128x64 .. 512x256 grids: every run of the ensemble equals the same run
  alone bit for bit after 20 steps, and two runs (one of them 512x256) equal the CPU oracle.

Bars: float64 to round-off (written per test); float32 within the tolerances written in each
test, which are the oracle's own float32-vs-float64 spread at the same step with a margin."""
from __future__ import annotations

import dataclasses

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
    a = a.detach().cpu().double().numpy() if torch.is_tensor(a) else np.asarray(a, dtype=np.float64)
    b = b.detach().cpu().double().numpy() if torch.is_tensor(b) else np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max())


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_default_flowsolver_matches_reference(golden, dtype):
    """FlowSolver(cfg) with the reference's defaults: the Lamb-Oseen start and 1, 10, 20 steps
    of the inlet comb's smoke, u, v, p, t and the coefficient history. float64 to 1e-11;
    float32: the start to 1e-6 (it passes through this build's own projection), the steps
    to 5e-5 (u, v, dye) and 5e-4 (p)."""
    _need_gpu()
    from aerolume_amd.ns2d.solver import FlowSolver
    f64 = dtype == "float64"
    F = golden(f"ns2d_default_{'f64' if f64 else 'f32'}.npz")
    fs = FlowSolver(cases.default_cfg(dtype, device="cuda"))
    st = fs.state
    e0 = max(_err(st.u, F["u_0"]), _err(st.v, F["v_0"]))
    print(f"default {dtype} start err {e0:.3e}", flush=True)
    assert e0 < (1e-13 if f64 else 1e-6)
    assert float(st.dye.abs().max()) == 0.0
    done = 0
    for n in (1, 10, 20):
        fs.advance(n - done)
        done = n
        st = fs.state
        eu = max(_err(st.u, F[f"u_{n}"]), _err(st.v, F[f"v_{n}"]))
        ed, ep = _err(st.dye, F[f"dye_{n}"]), _err(st.p, F[f"p_{n}"])
        print(f"default {dtype} step {n}: uv {eu:.3e} dye {ed:.3e} p {ep:.3e}", flush=True)
        assert eu < (1e-11 if f64 else 5e-5) and ed < (1e-11 if f64 else 5e-5), n
        assert ep < (1e-9 if f64 else 5e-4), n
        assert abs(float(st.t) - float(F[f"t_{n}"])) < (1e-13 if f64 else 1e-6)
    h = fs.history()
    for k in ("cd", "cl", "cm"):
        assert _err(h[k], F["hist:" + k]) < (1e-9 if f64 else 2e-3), k
    assert float(F["dye_20"].max()) > 0.5


@pytest.mark.parametrize("name", sorted(cases.VARIANTS2D))
def test_variants_match_reference(golden, name):
    """Five float64 steps of each 2D variant through step_full on the GPU: u, v to 1e-11 of the
    velocity scale, p to 1e-9, coefficients to 1e-9."""
    _need_gpu()
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d import solver as S
    F = golden("ns2d_variants.npz")
    cfg = cases.variant_cfg(name, device="cuda")
    ops = S.make_operators(cfg, device=DEV)
    geom = su.build_geometry(cfg.bodies, ops)
    t = lambda a: torch.as_tensor(a, device=DEV, dtype=ops.dtype)  # noqa: E731
    z = torch.zeros((), device=DEV, dtype=ops.dtype)
    cur = S.State(t(F[f"{name}:u0"]), t(F[f"{name}:v0"]), torch.zeros(ops.ny, ops.nx, device=DEV, dtype=ops.dtype),
                  None, z.clone(), z.clone())
    cs = []
    for _ in range(5):
        cur, c = S.step_full(cur, geom, ops)
        cs.append(c.cpu().numpy())
    scale = max(float(np.abs(F[f"{name}:u5"]).max()), 1.0)
    eu = max(_err(cur.u, F[f"{name}:u5"]), _err(cur.v, F[f"{name}:v5"]))
    ep = _err(cur.p, F[f"{name}:p5"])
    print(f"variant {name}: uv {eu:.3e} p {ep:.3e}", flush=True)
    assert eu < 1e-11 * scale
    assert ep < 1e-9 * max(1.0, float(np.abs(F[f"{name}:p5"]).max()))
    assert abs(float(cur.t) - float(F[f"{name}:t5"])) < 1e-14
    if cs[0].size:
        assert _err(np.stack(cs), F[f"{name}:coeffs"]) < 1e-9


def _c2_start(ops, seed=202):
    rng = np.random.default_rng(seed)
    u = 1.0 + 0.01 * rng.standard_normal((ops.ny, ops.nx + 1))
    v = 0.01 * rng.standard_normal((ops.ny + 1, ops.nx))
    return u, v


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_heated_airfoil_full_size_matches_oracle(dtype):
    """The 1024x512 heated airfoil (Smagorinsky, temperature body, Boussinesq, line emitter)
    through FlowSolver (graphs, chained steps, 1024- and 512-thread transform blocks) against
    the CPU oracle from the same projected noisy start at steps 1, 10 and 100: u, v, theta,
    dye, p and t. float64 to 1e-10 (fields of order 1) and 1e-8 (p). float32: 1e-5 on u, v,
    theta, dye and 5e-4 on p, four times the oracle's own float32-vs-float64 spread on this case
    over the same steps (u 1.4e-6, v 9.1e-7, theta 1.8e-6, dye 2.4e-6, p 1.3e-4; measured on
    the CPU with the oracle in both precisions from the same start)."""
    _need_gpu()
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d import solver as S
    from oracle import ns2d as O
    f64 = dtype == "float64"
    cfg_c = cases.c2_cfg(1024, 512, dtype)
    ops_c = su.make_operators(cfg_c, device="cpu")
    geom_c = su.build_geometry(cfg_c.bodies, ops_c, window=su.live_window(cfg_c.bodies, ops_c))
    u, v = _c2_start(ops_c)
    uc, vc = O.apply_boundary_values(torch.as_tensor(u, dtype=ops_c.dtype), torch.as_tensor(v, dtype=ops_c.dtype),
                                     ops_c)
    uc, vc, _, _ = O.project(uc, vc, ops_c)
    z = torch.zeros((), dtype=ops_c.dtype)
    zc = torch.zeros(ops_c.ny, ops_c.nx, dtype=ops_c.dtype)
    ref = O.State(uc, vc, zc.clone(), zc.clone(), z.clone(), z.clone(), zc.clone())
    fs = S.FlowSolver(cases.c2_cfg(1024, 512, dtype, device="cuda"))
    assert tuple(fs.geometry.window) == tuple(geom_c.window)
    fs.set_velocity(uc.to(DEV), vc.to(DEV))
    torch.set_flush_denormal(True)
    done = 0
    try:
        for n in (1, 10, 100):
            for _ in range(n - done):
                ref, _ = O.step_full(ref, geom_c, ops_c)
            fs.advance(n - done)
            done = n
            st = fs.state
            errs = {k: _err(getattr(st, k), getattr(ref, k)) for k in ("u", "v", "theta", "dye", "p")}
            print(f"c2 1024x512 {dtype} step {n}: " + " ".join(f"{k} {e:.3e}" for k, e in errs.items()), flush=True)
            for k in ("u", "v", "theta", "dye"):
                assert errs[k] < (1e-10 if f64 else 1e-5), (k, n)
            assert errs["p"] < (1e-8 if f64 else 5e-4), n
            assert abs(float(st.t) - float(ref.t)) < (1e-13 if f64 else 1e-6 * max(1.0, float(ref.t)))
    finally:
        torch.set_flush_denormal(False)
    div = float(S.divergence(fs.state.u, fs.state.v, fs.ops.h).abs().max())
    ref_div = float(O.divergence(ref.u, ref.v, ops_c.h).abs().max())
    rms = float(S.divergence(fs.state.u, fs.state.v, fs.ops.h).double().pow(2).mean().sqrt())
    ref_rms = float(O.divergence(ref.u, ref.v, ops_c.h).double().pow(2).mean().sqrt())
    print(f"c2 1024x512 {dtype} max|div| gpu {div:.3e} oracle {ref_div:.3e}; rms gpu {rms:.3e} oracle "
          f"{ref_rms:.3e}", flush=True)
    # post-projection divergence: the maximum under the rounding floor (half an ulp of u on each
    # of a cell's four faces, over h) or the oracle's maximum; the RMS within 2 % of the
    # oracle's (measured: float32 5.112e-6 against 5.108e-6, float64 9.551e-15 against 9.496e-15)
    eps = float(np.spacing(np.float64(1.0) if f64 else np.float32(1.0)))
    floor = 2.0 * eps * float(fs.state.u.abs().max()) / fs.ops.h
    assert div <= max(ref_div, floor)
    assert rms <= 1.02 * ref_rms


def test_sweep_64_runs_equal_single_runs_and_oracle():
    """The 64-run sweep (NACA 0012, 128x64 .. 512x256, alpha U[-5, 15], Re 10^U[3, 5], FP32)
    as one ensemble, 20 steps: every run bit for bit equal to the same run stepped alone; and
    two runs (the smallest and a 512x256 one) against the CPU oracle for 3 steps
    (float32: 2e-6 on u, v; 5e-5 on p)."""
    _need_gpu()
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d.ensemble import Ensemble2D, sweep_configs
    from aerolume_amd.ns2d.solver import FlowSolver
    from oracle import ns2d as O
    cfgs = sweep_configs(n=64, seed=404)
    sizes = [c.nx for c in cfgs]
    assert min(sizes) >= 128 and max(sizes) == 512
    ens = Ensemble2D(cfgs)
    starts = [(s.state.u.clone(), s.state.v.clone()) for s in ens.solvers]
    ens.advance(20)
    torch.cuda.synchronize()
    print("sweep: ensemble stepped", flush=True)
    for i, (c, se) in enumerate(zip(cfgs, ens.solvers)):
        alone = FlowSolver(dataclasses.replace(c, use_cuda_graphs=False), tile_rows=ens.tile_rows)
        alone.advance(20)
        torch.cuda.synchronize()
        for f in ("u", "v", "p", "t", "dt"):
            assert torch.equal(getattr(se.state, f), getattr(alone.state, f)), (i, c.nx, f)
        del alone
    # oracle on the smallest run and on a 512x256 one, 3 steps from the runs' own start
    picks = [int(np.argmin(sizes)), sizes.index(512)]
    ens2 = Ensemble2D([cfgs[i] for i in picks])
    ens2.advance(3)
    torch.cuda.synchronize()
    for k, i in enumerate(picks):
        c = dataclasses.replace(cfgs[i], device="cpu")
        ops = su.make_operators(c, device="cpu")
        geom = su.build_geometry(c.bodies, ops, window=su.live_window(c.bodies, ops))
        z = torch.zeros((), dtype=ops.dtype)
        u0, v0 = (a.cpu() for a in starts[i])
        cur = O.State(u0, v0, torch.zeros(ops.ny, ops.nx, dtype=ops.dtype), None, z.clone(), z.clone())
        for _ in range(3):
            cur, _ = O.step_full(cur, geom, ops)
        st = ens2.solvers[k].state
        eu, ep = max(_err(st.u, cur.u), _err(st.v, cur.v)), _err(st.p, cur.p)
        print(f"sweep run {i} ({c.nx}x{c.ny}) vs oracle: uv {eu:.3e} p {ep:.3e}", flush=True)
        assert eu < 2e-6 and ep < 5e-5


@pytest.mark.parametrize("name", ["ww_tunnel_moving", "ww_closed_room"])
def test_wall_model_float32_matches_oracle(golden, name):
    """The wall-model variants in float32 (whose power law runs on the hardware base-2
    logarithm and exponential) against the float32 oracle (library powers) from the same
    start, five steps: u, v within 2e-6 of the velocity scale."""
    _need_gpu()
    import oracle.ns2d as O
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d import solver as S
    F = golden("ns2d_variants.npz")
    cfg = cases.variant_cfg(name, dtype="float32", device="cuda")
    ops = S.make_operators(cfg, device=DEV)
    geom = su.build_geometry(cfg.bodies, ops)
    cfg_c = cases.variant_cfg(name, dtype="float32")
    ops_c = su.make_operators(cfg_c, device="cpu")
    geom_c = su.build_geometry(cfg_c.bodies, ops_c)
    t = lambda a, d: torch.as_tensor(a, device=d, dtype=torch.float32)  # noqa: E731
    z = lambda d: torch.zeros((), device=d, dtype=torch.float32)  # noqa: E731
    cur = S.State(t(F[f"{name}:u0"], DEV), t(F[f"{name}:v0"], DEV), torch.zeros(ops.ny, ops.nx, device=DEV), None,
                  z(DEV), z(DEV))
    ref = O.State(t(F[f"{name}:u0"], "cpu"), t(F[f"{name}:v0"], "cpu"), torch.zeros(ops.ny, ops.nx), None, z("cpu"),
                  z("cpu"))
    for _ in range(5):
        cur, _ = S.step_full(cur, geom, ops)
        ref, _ = O.step_full(ref, geom_c, ops_c)
    scale = max(float(ref.u.abs().max()), 1.0)
    eu = max(_err(cur.u, ref.u.numpy()), _err(cur.v, ref.v.numpy()))
    print(f"wall model float32 {name}: {eu:.3e}", flush=True)
    assert eu < 2e-6 * scale
