"""Benchmark of the MI355X Navier-Stokes step (one JSON line on stdout).

Workload at N=1: the 2D heated-airfoil case at 1024x512 cells, FP32 - NACA 0012 at 8 deg held
at theta = 1 (dT 10 K), Re 1e5, Smagorinsky Cs 0.17, upwind3 + SSP-RK3 at CFL 0.5, exact DCT
projection, Boussinesq coupling (Pr 0.71, Ri 0.1), one passive smoke scalar from a line
emitter; seeded 1 % Gaussian velocity noise on the inflow state (synthetic data, no files).
A step is one full SSP-RK3 time step of that case through ``FlowSolver.advance`` (HIP graphs
of 10 steps). ``value`` is Mcell-steps/s = 524288 cells x steps / s / 1e6.

Also reported: ``roofline`` of the dominant kernel (the fused stage kernel; algorithmic bytes
per launch / its average duration timed with HIP events on the solver's stream) and
``cpu_baseline`` (the CPU oracle on the host's cores, a bounded sample of the same case).

``workloads.ensemble``: config 4, a 64-run angle / Reynolds sweep on uneven grids as one
batched ensemble (solves/s for 1000-step solves) against the same solves run one after
another on the same GPU.

``workloads.closure_cnn``: the learned closure's network (9 -> 32 -> 32 -> 32 -> 32 -> 1 channels
of 3x3x3 periodic convolutions, bf16 weights, float32 accumulation) on a 64^3 grid: inference,
training forward and backward in ms (best of three blocks of 10 calls, HIP events), and the
matrix throughput of the forward pass against the 2.5 PFLOP/s dense bf16 peak.

``workloads.3d_car`` carries the same figures for the 3D car case at 256x128x128 (4.19M
cells, FP32): a 4 x 1.6 x 1.2 rounded box on four rolling wheels over a floor moving at U,
Re 5e4, WALE, periodic in y, upwind3 + SSP-RK3 at CFL 0.5, exact DCT/DHT projection;
seeded 1 % Gaussian velocity noise on the inflow state; ``FlowSolver3D.advance``.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "Mcell-steps/s at 2D 1024×512 & 3D 4.2M cells; %HBM peak; ensemble solves/s"
HBM_PEAK_GBS = 8000.0
NX, NY = 1024, 512
SEED = 202
# algorithmic bytes per cell of one stage-kernel launch at stages 2-3 (FP32): read u, v, nu_t,
# theta of the stage input and u0, v0, theta0 of the step start (7 x 4 B), write u, v, theta
# of the stage (3 x 4 B)
STAGE_BYTES_PER_CELL = 40
STEP_BYTES_PER_CELL = 300     # whole fused step of this case (byte model of the design notes)
NX3, NY3, NZ3 = 256, 128, 128
SEED3 = 303
# 3D stage kernel at stages 2-3 (FP32): read u^, v^, w^ of the previous stage, psi, and u0,
# v0, w0 of the step start (7 x 4 B), write u^, v^, w^ (3 x 4 B)
STAGE3_BYTES_PER_CELL = 40
# whole step: first stage 28 + two stages 40 + three projections (x pass 16, the fused (z, y)
# plane pass 8 + its 1/lambda table 4, inverse x pass 16) + final 28 + nu_t saved by the final
# kernel and read by the next first stage 8
STEP3_BYTES_PER_CELL = 276
# HBM bytes per stage-2 launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by the
# same counters on a 512 MiB device copy (fetch x2.000, write x1.000):
# profiles/stage_kernels_hbm_counters.md
TRAFFIC_STAGE_2D = 28.87e6
TRAFFIC_STAGE_3D = 183.90e6


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


#: graph blocks every leg warms up before its clock starts: each HIP graph a timed block replays
#: is captured (and replayed once) before the first timed step
WARM_BLOCKS = 2
#: timed blocks per leg; the figure reported is the best block (the reference's protocol,
#: validation/bench_ns2d.py:71-90, validation/bench_ns3d.py:73-101)
TIMED_BLOCKS = 3


def warm_steps(warmup: int, graph_steps: int = 10) -> int:
    """Untimed steps before the clock: at least ``WARM_BLOCKS`` whole graph blocks, so that no
    timed step pays for a graph capture, whatever ``--warmup`` says."""
    n = max(int(warmup), WARM_BLOCKS * graph_steps)
    return -(-n // graph_steps) * graph_steps


def timed_blocks(advance, steps: int, blocks: int = TIMED_BLOCKS) -> list:
    """Wall seconds of ``blocks`` runs of exactly ``steps`` steps of ``advance``, each bracketed
    by a device synchronisation on both sides."""
    import torch
    walls = []
    for _ in range(blocks):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        advance(steps)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    return walls


def host_cpu() -> dict:
    """Model name and physical core count of this host (from /proc/cpuinfo), and the CPU share
    this process may use (its affinity mask)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def c2_config(dtype="float32", device="cuda"):
    import math

    import numpy as np

    from aerolume_amd.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig
    t = 0.12
    beta = np.linspace(0.0, np.pi, 201)
    x = 0.5 * (1 - np.cos(beta))
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 - 0.1036 * x**4)
    foil = np.vstack([np.c_[x, yt][::-1], np.c_[x, -yt][1:-1]])
    a = math.radians(-8.0)
    c, s = math.cos(a), math.sin(a)
    p = foil - (0.25, 0.0)
    foil = np.c_[c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1]] + (0.25, 0.0)
    return SolverConfig(nx=NX, ny=NY, h=8.0 / NX, x0=-2.0, y0=-2.0, re=1e5,
                        bodies=[ImmersedBody(foil, thermal="temperature", thermal_value=1.0)],
                        advection="upwind3", turbulence="smagorinsky", cfl=0.5, dtype=dtype, device=device,
                        dye=True, dye_inlet=False, perturb=False,
                        thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                        emitters=[EmitterSpec(kind="line", position=(-1.0, -0.2), end=(-1.0, 0.2), channel=0,
                                              rate=1.0)])


def noisy_velocity(ops, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    u = 1.0 + 0.01 * rng.standard_normal((ops.ny, ops.nx + 1))
    v = 0.01 * rng.standard_normal((ops.ny + 1, ops.nx))
    return u, v


def gpu_run(steps: int, warmup: int):
    import torch

    from aerolume_amd.ns2d import solver as S
    from aerolume_amd import _lib
    cfg = c2_config()
    solver = S.FlowSolver(cfg)
    ops = solver.ops
    u, v = noisy_velocity(ops, SEED)
    u = torch.as_tensor(u, device=ops.device, dtype=ops.dtype)
    v = torch.as_tensor(v, device=ops.device, dtype=ops.dtype)
    u, v = S.apply_boundary_values(u, v, ops)
    u, v, _ = S.project(u, v, ops)
    solver.set_velocity(u, v)
    log(f"bench: workspace {solver.workspace_bytes / 2**20:.1f} MiB, window {solver.geometry.window}")
    solver.advance(warm_steps(warmup))
    walls = timed_blocks(solver.advance, steps)
    log(f"bench: 2D blocks of {steps} steps: " + ", ".join(f"{w * 1e3:.3f} ms" for w in walls))
    if not all(bool(torch.isfinite(t).all()) for t in (solver.state.u, solver.state.v, solver.state.theta)):
        raise RuntimeError("non-finite state after the timed steps")
    # dominant kernel: the stage kernel, timed alone with HIP events on the solver's stream
    lib = _lib.load()
    ms = ctypes.c_double()
    pl = solver._plan
    iters = 200
    st = _lib.stream_handle(ops.device)
    rc = lib.aero2d_time_stage(pl.handle, ctypes.c_void_p(st), ctypes.byref(solver._fields[solver._parity][0]),
                               ctypes.byref(pl.geom_ptrs), ctypes.byref(pl.params), 1, iters, ctypes.byref(ms))
    _lib.check(rc, "aero2d_time_stage")
    torch.cuda.synchronize()
    return solver, walls, ms.value


def gpu_run_3d(steps: int, warmup: int):
    import torch

    from aerolume_amd.ns3d import solver as S
    from tests.cases import c3_cfg
    cfg = c3_cfg("float32", device="cuda")
    solver = S.FlowSolver3D(cfg)
    ops = solver.ops
    u, v, w = noisy_velocity3d(ops, SEED3)
    dev = ops.device
    u, v, w = S.apply_boundary_values(torch.as_tensor(u, device=dev, dtype=ops.dtype),
                                      torch.as_tensor(v, device=dev, dtype=ops.dtype),
                                      torch.as_tensor(w, device=dev, dtype=ops.dtype), ops)
    u, v, w, _ = S.project3d(u.contiguous(), v.contiguous(), w.contiguous(), ops)
    solver.set_velocity(u, v, w)
    log(f"bench3d: workspace {solver.workspace_bytes / 2**20:.1f} MiB, window {solver.geometry.window}")
    solver.advance(warm_steps(warmup))
    walls = timed_blocks(solver.advance, steps)
    log(f"bench3d: blocks of {steps} steps: " + ", ".join(f"{w * 1e3:.3f} ms" for w in walls))
    st = solver.state
    if not all(bool(torch.isfinite(t).all()) for t in (st.u, st.v, st.w, st.p)):
        raise RuntimeError("non-finite 3D state after the timed steps")
    stage_ms = solver._plan.time_stage(solver._fields, 1, 100)
    torch.cuda.synchronize()
    return solver, walls, stage_ms


def gpu_run_ensemble(steps: int, warmup: int, n: int = 64, seed: int = 404):
    """64-run sweep (NACA 0012, uneven grids 128..512 x 64..256, alpha U[-5, 15] deg,
    Re 10^U[3, 5], FP32, no LES): the batched ensemble against the same runs solved one after
    another with FlowSolver (HIP graphs on both sides). The ensemble is timed with the default
    launch shapes (8-row tiles, 2 column slots), then with the shapes its autotuner picks;
    ``value`` is the tuned rate."""
    import torch

    from aerolume_amd.ns2d.ensemble import Ensemble2D, sweep_configs
    from aerolume_amd.ns2d.solver import FlowSolver
    cfgs = sweep_configs(n=n, seed=seed)
    ens = Ensemble2D(cfgs)
    cells = ens.cells

    def timed():
        ens.advance(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.advance(steps)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    wall_default = timed()
    t0 = time.perf_counter()
    choice = ens.tune(iters=5)
    tune_s = time.perf_counter() - t0
    wall_e = timed()
    ok = all(bool(torch.isfinite(s.state.u).all()) for s in ens.solvers)
    del ens
    torch.cuda.empty_cache()
    wall_s = 0.0
    for c in cfgs:
        fs = FlowSolver(c)
        fs.advance(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fs.advance(steps)
        torch.cuda.synchronize()
        wall_s += time.perf_counter() - t0
        del fs
    if not ok:
        raise RuntimeError("non-finite ensemble state")
    return {"value": round(n / wall_e, 3), "unit": "solves/s", "sequential_value": round(n / wall_s, 3),
            "speedup": round(wall_s / wall_e, 2), "untuned_value": round(n / wall_default, 3),
            "tuned": dict(choice, seconds=round(tune_s, 3)), "runs": n, "steps_per_solve": steps, "warmup": warmup,
            "cells_total": cells, "mcell_steps_per_s": round(cells * steps / wall_e / 1e6, 1),
            "sequential_mcell_steps_per_s": round(cells * steps / wall_s / 1e6, 1),
            "ms_per_ensemble_step": round(wall_e / steps * 1e3, 4), "dtype": "f32", "data": "synthetic",
            "config": {"workload": "64-run NACA 0012 sweep, nx in 5-smooth [128, 512], ny = nx/2, h = 8/nx, "
                                   "alpha U[-5, 15] deg, Re 10^U[3, 5], upwind3, no LES, FP32", "seed": seed}}


def _rel_err(a, b) -> dict:
    d = (a.double() - b.double()).abs()
    scale = float(b.double().abs().max())
    return {"max_rel": float(d.max()) / max(scale, 1e-300),
            "l2_rel": float(d.norm()) / max(float(b.double().norm()), 1e-300)}


def _fp64_report(s32, s64, fields, steps) -> dict:
    out = {"steps": steps, "t_fp64": float(s64.state.t), "t_rel": abs(float(s32.state.t) / float(s64.state.t) - 1.0)}
    for f in fields:
        err = _rel_err(getattr(s32.state, f), getattr(s64.state, f))
        out[f] = {k: float(f"{v:.3e}") for k, v in err.items()}
    out["t_rel"] = float(f"{out['t_rel']:.3e}")
    return out


def fp64_check_2d(steps: int = 200) -> dict:
    """The 2D workload in FP32 against the same kernels in FP64 from the same start (the FP64
    projected seed, rounded for the FP32 run), ``steps`` steps each on its own CFL dt: the FP32
    result's relative error (max and L2 over the field) and its time drift."""
    import torch

    from aerolume_amd.ns2d import solver as S
    s64 = S.FlowSolver(c2_config(dtype="float64"))
    ops = s64.ops
    u, v = noisy_velocity(ops, SEED)
    u, v = S.apply_boundary_values(torch.as_tensor(u, device=ops.device, dtype=ops.dtype),
                                   torch.as_tensor(v, device=ops.device, dtype=ops.dtype), ops)
    u, v, _ = S.project(u, v, ops)
    s64.set_velocity(u, v)
    s32 = S.FlowSolver(c2_config(dtype="float32"))
    s32.set_velocity(u.float(), v.float())
    s64.advance(steps)
    s32.advance(steps)
    torch.cuda.synchronize()
    out = _fp64_report(s32, s64, ("u", "v", "p", "theta"), steps)
    c32, c64 = s32.coefficients(), s64.coefficients()
    out["coefficients_max_abs"] = float(f"{float(abs(c32 - c64).max()):.3e}")
    return out


def fp64_check_3d(steps: int = 100) -> dict:
    """The 3D car workload in FP32 against the same kernels in FP64 from the same start."""
    import torch

    from aerolume_amd.ns3d import solver as S
    from tests.cases import c3_cfg
    s64 = S.FlowSolver3D(c3_cfg("float64", device="cuda"))
    ops = s64.ops
    dev = ops.device
    u, v, w = (torch.as_tensor(a, device=dev, dtype=ops.dtype) for a in noisy_velocity3d(ops, SEED3))
    u, v, w = S.apply_boundary_values(u, v, w, ops)
    u, v, w, _ = S.project3d(u.contiguous(), v.contiguous(), w.contiguous(), ops)
    s64.set_velocity(u, v, w)
    s32 = S.FlowSolver3D(c3_cfg("float32", device="cuda"))
    s32.set_velocity(u.float(), v.float(), w.float())
    s64.advance(steps)
    s32.advance(steps)
    torch.cuda.synchronize()
    return _fp64_report(s32, s64, ("u", "v", "w", "p"), steps)


def noisy_velocity3d(ops, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    u = 1.0 + 0.01 * rng.standard_normal((ops.nz, ops.ny, ops.nx + 1))
    v = 0.01 * rng.standard_normal((ops.nz, ops.ny + 1, ops.nx))
    w = 0.01 * rng.standard_normal((ops.nz + 1, ops.ny, ops.nx))
    return u, v, w


def cpu_blocks(step, threads: int, block_steps: int, blocks: int = TIMED_BLOCKS) -> float:
    """Best wall seconds per step over ``blocks`` blocks of ``block_steps`` calls of ``step``
    on ``threads`` torch CPU threads with flush-to-zero (the reference's CPU stepping mode,
    solvers/fpmode.py:45-63), after one warm-up call."""
    import torch
    old = torch.get_num_threads()
    torch.set_num_threads(threads)
    torch.set_flush_denormal(True)
    try:
        step()
        best = float("inf")
        for _ in range(blocks):
            t0 = time.perf_counter()
            for _ in range(block_steps):
                step()
            best = min(best, (time.perf_counter() - t0) / block_steps)
    finally:
        torch.set_flush_denormal(False)
        torch.set_num_threads(old)
    return best


def cpu_report(step, cells: int, block_steps: int, what: str) -> dict:
    """``cpu_baseline`` of one case: the CPU oracle timed at all the threads torch uses here
    (``value``) and at 4 threads (the reference's published CPU protocol,
    docs/theory_ns2d.md:426-428), best of three blocks each."""
    import torch
    threads = torch.get_num_threads()
    host = host_cpu()
    t_all = cpu_blocks(step, threads, block_steps)
    t_4 = cpu_blocks(step, 4, block_steps)
    return {"value": round(cells / t_all / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "value_4_threads": round(cells / t_4 / 1e6, 3), "host": host,
            "sample": f"{what}, FP32, torch CPU oracle with flush-to-zero: one warm-up step, then the best of "
                      f"{TIMED_BLOCKS} blocks of {block_steps} steps, at {threads} threads (value) and at 4 "
                      f"threads (value_4_threads); host {host['model']}, {host['physical_cores']} physical "
                      f"cores, {host['affinity_cpus']} in this process's CPU share"}


def cpu_baseline_3d(block_steps: int = 3):
    """The CPU oracle on the 3D car case (see ``cpu_report``)."""
    import torch

    from aerolume_amd.ns3d import setup as su
    from oracle import ns3d as O
    from tests.cases import c3_cfg
    cfg = c3_cfg("float32", device="cpu")
    ops = su.make_operators3d(cfg, device="cpu")
    geom = su.build_geometry3d(cfg.bodies, ops)
    u, v, w = (torch.as_tensor(a, dtype=ops.dtype) for a in noisy_velocity3d(ops, SEED3))
    u, v, w = O.apply_boundary_values(u, v, w, ops)
    u, v, w, _ = O.project3d(u, v, w, ops)
    z = torch.zeros((), dtype=ops.dtype)
    box = [O.State3D(u, v, w, torch.zeros(ops.nz, ops.ny, ops.nx, dtype=ops.dtype), z.clone(), z.clone())]

    def step():
        box[0], _ = O.step3d_full(box[0], geom, ops)
    return cpu_report(step, NX3 * NY3 * NZ3, block_steps, "the 256x128x128 car case")


def cpu_baseline(block_steps: int = 60):
    """The CPU oracle on the 2D heated-airfoil case (see ``cpu_report``); 60-step blocks as
    validation/bench_ns2d.py:81."""
    import torch

    from aerolume_amd.ns2d import setup as su
    from oracle import ns2d as O
    cfg = c2_config(device="cpu")
    ops = su.make_operators(cfg, device="cpu")
    geom = su.build_geometry(cfg.bodies, ops, window=su.live_window(cfg.bodies, ops))
    u, v = noisy_velocity(ops, SEED)
    u, v = O.apply_boundary_values(torch.as_tensor(u, dtype=ops.dtype), torch.as_tensor(v, dtype=ops.dtype), ops)
    u, v, _, _ = O.project(u, v, ops)
    z = torch.zeros((), dtype=ops.dtype)
    box = [O.State(u, v, torch.zeros(ops.ny, ops.nx, dtype=ops.dtype), torch.zeros(ops.ny, ops.nx, dtype=ops.dtype),
                   z.clone(), z.clone(), torch.zeros(ops.ny, ops.nx, dtype=ops.dtype))]

    def step():
        box[0], _ = O.step_full(box[0], geom, ops)
    return cpu_report(step, NX * NY, block_steps, "the 1024x512 heated-airfoil case")


CNN_GRID = 64
CNN_CHANNELS = (9, 32, 32, 32, 32, 1)
BF16_PEAK_TFLOPS = 2500.0


def gpu_run_cnn(grid: int = CNN_GRID, calls: int = 10) -> dict:
    """The closure network at grid^3: ms per inference / training forward / backward."""
    import torch
    from aerolume_amd.closure.network import CNN3D
    net = CNN3D((grid, grid, grid), CNN_CHANNELS, seed=1)
    g = torch.Generator(device="cuda").manual_seed(505)
    x = torch.randn(CNN_CHANNELS[0], grid, grid, grid, device="cuda", generator=g)
    dy = torch.randn(CNN_CHANNELS[-1], grid, grid, grid, device="cuda", generator=g)
    y = torch.empty(CNN_CHANNELS[-1], grid, grid, grid, device="cuda")

    def best(fn) -> float:
        fn()
        torch.cuda.synchronize()
        out = []
        for _ in range(TIMED_BLOCKS):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(calls):
                fn()
            e1.record()
            torch.cuda.synchronize()
            out.append(e0.elapsed_time(e1) / calls)
        return min(out)

    infer = best(lambda: net.infer(x, out=y))
    fwd = best(lambda: net.forward(x))
    net.forward(x)
    bwd = best(lambda: net.backward(dy))
    cells = grid ** 3
    # useful flops of the forward pass: 2 x 27 x cin x cout per cell per layer
    flops = sum(2 * 27 * ci * co for ci, co in zip(CNN_CHANNELS[:-1], CNN_CHANNELS[1:])) * cells
    return {"grid": [grid] * 3, "channels": list(CNN_CHANNELS), "infer_ms": round(infer, 4),
            "forward_ms": round(fwd, 4), "backward_ms": round(bwd, 4),
            "forward_gflop": round(flops / 1e9, 2), "infer_tflops": round(flops / (infer * 1e-3) / 1e12, 1),
            "infer_frac_of_bf16_peak": round(flops / (infer * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
            "dtype": "bf16 weights, f32 data", "data": "synthetic"}


def run_only(args) -> int:
    """One workload alone, for profiler runs: prints a short JSON summary."""
    import torch
    out = {"only": args.only}
    # at least two 10-step blocks of warm-up, so that every HIP graph the timed steps replay
    # is captured before the clock starts
    if args.only == "2d":
        _, walls, stage_ms = gpu_run(args.steps, args.warmup)
        out.update(ms_per_step=min(walls) / args.steps * 1e3, stage_us=stage_ms * 1e3)
    elif args.only == "3d":
        _, walls, stage_ms = gpu_run_3d(args.steps3d, args.warmup3d)
        out.update(ms_per_step=min(walls) / args.steps3d * 1e3, stage_us=stage_ms * 1e3)
    elif args.only == "ensemble":
        out.update(gpu_run_ensemble(args.ensemble_steps, 20))
    elif args.only == "cnn":
        out.update(gpu_run_cnn())
    else:
        n = 512 * 2**20 // 4
        x = torch.rand(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(10):
            y.copy_(x)
        torch.cuda.synchronize()
        out.update(copy_bytes_read=n * 4, copy_bytes_written=n * 4, launches=10)
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--steps3d", type=int, default=500)
    ap.add_argument("--warmup3d", type=int, default=20)
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D car workload")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--only", choices=("2d", "3d", "ensemble", "cnn", "copy"), default=None,
                    help="run one workload and print its figures (profiling runs); 'copy' times a plain "
                         "512 MiB device copy (the HBM counter calibration)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble workload")
    ap.add_argument("--no-cnn", action="store_true", help="skip the closure-network workload")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box has one "
              "MI355X); run with --gpus 1", file=sys.stderr)
        return 2
    if args.only is not None:
        return run_only(args)
    solver, walls, stage_ms = gpu_run(args.steps, args.warmup)
    wall = min(walls)
    cells = NX * NY
    value = cells * args.steps / wall / 1e6
    ms_step = wall / args.steps * 1e3
    stage_bytes = STAGE_BYTES_PER_CELL * cells
    achieved = stage_bytes / (stage_ms * 1e-3) / 1e9
    step_gbs = STEP_BYTES_PER_CELL * cells / (wall / args.steps) / 1e9
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": warm_steps(args.warmup), "ms_per_step": round(ms_step, 5),
        "timed_blocks_ms": [round(w * 1e3, 4) for w in walls], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D 1024x512 heated NACA 0012 at 8 deg, Re 1e5, Smagorinsky, Boussinesq Pr 0.71 "
                               "Ri 0.1, line smoke emitter, upwind3 + SSP-RK3 CFL 0.5, DCT projection",
                   "nx": NX, "ny": NY, "cells": cells, "seed": SEED, "graph_steps": 10},
        "roofline": {"bound": "hbm", "kernel": "k_stage (stage 2 of 3)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_STAGE_2D, "avg_launch_us": round(stage_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": stage_bytes},
        "step_hbm": {"algorithmic_bytes_per_step": STEP_BYTES_PER_CELL * cells, "achieved_gbs": round(step_gbs, 1),
                     "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    del solver
    line["fp64_check"] = fp64_check_2d()
    if not args.no_3d:
        s3, walls3, stage3_ms = gpu_run_3d(args.steps3d, args.warmup3d)
        wall3 = min(walls3)
        cells3 = NX3 * NY3 * NZ3
        v3 = cells3 * args.steps3d / wall3 / 1e6
        sb3 = STAGE3_BYTES_PER_CELL * cells3
        ach3 = sb3 / (stage3_ms * 1e-3) / 1e9
        step3 = STEP3_BYTES_PER_CELL * cells3 / (wall3 / args.steps3d) / 1e9
        w3 = {"value": round(v3, 2), "unit": "Mcell-steps/s", "steps": args.steps3d,
              "warmup": warm_steps(args.warmup3d), "ms_per_step": round(wall3 / args.steps3d * 1e3, 5),
              "timed_blocks_ms": [round(w * 1e3, 4) for w in walls3], "dtype": "f32", "data": "synthetic",
              "config": {"workload": "3D 256x128x128 car (4 x 1.6 x 1.2 rounded box on 4 rolling wheels), "
                                     "moving floor, periodic y, Re 5e4, WALE, upwind3 + SSP-RK3 CFL 0.5, "
                                     "DCT/DHT projection", "nx": NX3, "ny": NY3, "nz": NZ3, "cells": cells3,
                         "seed": SEED3, "graph_steps": 10},
              "roofline": {"bound": "hbm", "kernel": "k_stage 3D (stage 2 of 3)", "achieved": round(ach3, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach3 / HBM_PEAK_GBS, 4),
                           "traffic": TRAFFIC_STAGE_3D, "avg_launch_us": round(stage3_ms * 1e3, 3),
                           "algorithmic_bytes_per_launch": sb3},
              "step_hbm": {"algorithmic_bytes_per_step": STEP3_BYTES_PER_CELL * cells3,
                           "achieved_gbs": round(step3, 1), "frac": round(step3 / HBM_PEAK_GBS, 4)}}
        del s3
        w3["fp64_check"] = fp64_check_3d()
        if not args.no_cpu:
            w3["cpu_baseline"] = cpu_baseline_3d()
            w3["gpu_over_cpu"] = round(v3 / w3["cpu_baseline"]["value"], 1)
        line.setdefault("workloads", {})["3d_car"] = w3
    if not args.no_ensemble:
        line.setdefault("workloads", {})["ensemble"] = gpu_run_ensemble(args.ensemble_steps, 20)
    if not args.no_cnn:
        line.setdefault("workloads", {})["closure_cnn"] = gpu_run_cnn()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
