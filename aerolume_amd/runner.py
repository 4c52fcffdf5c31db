"""Run loops around the solvers: chunked advancing with the reference runner's health checks,
and factorial sweeps routed to the batched 2D ensemble.

``run_chunked`` follows the reference's ``run_cfd`` stepping rules (simulation/runner.py:444-600):
a short first chunk (it pays for the graph capture), then chunks of whole graph blocks; after
each chunk the run stops with :class:`FlowDiverged` when a force coefficient is not finite or
exceeds ``COEFF_LIMIT``, or when the CFL-limited time step has collapsed below ``DT_COLLAPSE``
of the planned one (the velocity grows without bound long before the forces turn NaN); running
means start at ``t_avg``. ``sweep_points`` builds the full factorial of list-valued parameters in
the order of the reference's ``Project.sweep`` (api.py:4877-4930), and ``run_sweep`` advances a
list of 2D configs in :class:`~aerolume_amd.ns2d.ensemble.Ensemble2D` batches (one launch per
kernel for all runs) instead of one solve after another, with per-run results as the sequential
solves give them.
"""
from __future__ import annotations

import itertools
import math
import time
from typing import Callable

import numpy as np

#: A time step below this fraction of the planned one means the flow is blowing up
#: (reference simulation/runner.py DT_COLLAPSE).
DT_COLLAPSE = 1e-3
#: Steps of the first chunk: two graph blocks (reference FIRST_CHUNK_STEPS).
FIRST_CHUNK_STEPS = 20
#: Force coefficients beyond this magnitude count as a blow-up.
COEFF_LIMIT = 1e6


class FlowDiverged(RuntimeError):
    """The flow blew up: non-finite or huge coefficients, or a collapsed time step."""


def check_health(solver, planned_dt: float) -> None:
    """Raise :class:`FlowDiverged` when the solver's latest state has blown up (reference
    runner.py:554-559)."""
    t = float(solver.t)
    c = np.asarray(solver.coefficients(), dtype=np.float64)
    if c.size and (not np.all(np.isfinite(c)) or np.abs(c).max() > COEFF_LIMIT):
        raise FlowDiverged(f"the flow diverged at t = {t:.4g}: force coefficients {c.tolist()}")
    dt = float(solver.dt)
    if planned_dt > 0 and not dt > DT_COLLAPSE * planned_dt:
        raise FlowDiverged(f"the flow diverged at t = {t:.4g}: the time step collapsed to {dt:.3g} "
                           f"(planned {planned_dt:.3g}); the velocity is growing without bound")


def chunk_sizes(t: float, target: float, dt: float, budget_steps: int, first: bool, graph_steps: int) -> int:
    """Steps of the next chunk: at most ``budget_steps`` (the first chunk at most
    FIRST_CHUNK_STEPS), no further than ``target`` at the current dt, whole graph blocks from
    two blocks on (reference runner.py:512-527)."""
    n = max(1, int(budget_steps))
    if first:
        n = min(n, FIRST_CHUNK_STEPS)
    if dt > 0:
        n = min(n, max(1, math.ceil((target - t) / dt)))
    if n >= 2 * graph_steps:
        n -= n % graph_steps
    return n


def run_chunked(solver, t_end: float, *, planned_dt: float | None = None, t_avg: float | None = None,
                chunk_steps: int = 200, graph_steps: int = 10, progress: Callable | None = None,
                cancel=None) -> dict:
    """Advance ``solver`` (2D or 3D) to ``t_end`` in chunks with the health checks after each
    chunk; ``t_avg`` starts the running means there. ``progress(fraction, t)`` is called per
    chunk, ``cancel`` (an object with ``is_set()``) stops between chunks. Returns the time
    reached, steps, chunks, stepping wall time and rate, whether it was cancelled and when
    averaging started."""
    planned = float(planned_dt) if planned_dt is not None else float(solver.dt)
    t0, steps0 = float(solver.t), int(solver.steps)
    averaging_from = None
    first, chunks, wall = True, 0, 0.0
    cancelled = False
    while float(solver.t) < t_end - 1e-12:
        if cancel is not None and cancel.is_set():
            cancelled = True
            break
        t = float(solver.t)
        averaging = averaging_from is not None or t_avg is None
        target = t_end if averaging else min(t_end, t_avg)
        n = chunk_sizes(t, target, float(solver.dt), chunk_steps, first, graph_steps)
        w0 = time.perf_counter()
        solver.advance(n)
        wall += time.perf_counter() - w0
        first = False
        chunks += 1
        check_health(solver, planned)
        if t_avg is not None and averaging_from is None and float(solver.t) >= t_avg - 1e-12:
            solver.start_averaging()
            averaging_from = float(solver.t)
        if progress is not None:
            progress(min(1.0, (float(solver.t) - t0) / max(t_end - t0, 1e-300)), float(solver.t))
    steps = int(solver.steps) - steps0
    return {"t": float(solver.t), "steps": steps, "chunks": chunks, "wall_s": wall,
            "steps_per_s": steps / wall if wall > 0 else float("nan"), "cancelled": cancelled,
            "averaging_from": averaging_from}


def sweep_points(**params) -> list[dict]:
    """The full factorial of the list-valued parameters (scalars are passed to every point), in
    the order of the reference's ``Project.sweep``: the first swept name varies slowest."""
    swept, fixed = {}, {}
    for key, val in params.items():
        if isinstance(val, (list, tuple, range, np.ndarray)) and not isinstance(val, str):
            vals = [x.item() if isinstance(x, np.generic) else x for x in np.asarray(val, dtype=object).ravel()]
            if not vals:
                raise ValueError(f"{key}= is empty")
            swept[key] = vals
        else:
            fixed[key] = val
    if not swept:
        raise ValueError("nothing to sweep: give at least one parameter as a list")
    names = list(swept)
    return [dict(fixed, **dict(zip(names, combo))) for combo in itertools.product(*(swept[n] for n in names))]


def _batchable(cfg) -> bool:
    """Configs the batched step carries: no smoke, no heat transfer, no PID step, no wall model."""
    return (not cfg.dye and cfg.thermal is None and getattr(cfg, "adaptive_dt", "cfl") != "pid"
            and getattr(cfg, "wall_model", "none") == "none")


def run_sweep(configs: list, steps: int, *, average_last: float = 0.5, tune: bool = True,
              max_batch: int = 128) -> list[dict]:
    """Advance every 2D config by ``steps`` steps: configs the batched step carries go through
    :class:`~aerolume_amd.ns2d.ensemble.Ensemble2D` (batches of the same dtype, at most
    ``max_batch`` runs, autotuned launch shapes when ``tune``), the others one by one. Each
    result holds the run's final coefficients, their means and standard deviations over the last
    ``average_last`` share of the steps, t and the history; the order is that of ``configs``."""
    from .ns2d.ensemble import Ensemble2D
    from .ns2d.solver import FlowSolver

    def summary(solver) -> dict:
        h = solver.history()
        k0 = int(len(h["t"]) * (1.0 - average_last))
        tail = {c: np.asarray(h[c][k0:]) for c in ("cd", "cl", "cm")}
        return {"t": float(solver.t), "steps": int(solver.steps), "coefficients": solver.coefficients(),
                "mean": {c: float(v.mean()) if v.size else float("nan") for c, v in tail.items()},
                "std": {c: float(v.std()) if v.size else float("nan") for c, v in tail.items()},
                "history": h}

    results: list[dict | None] = [None] * len(configs)
    groups: dict = {}
    for i, cfg in enumerate(configs):
        if _batchable(cfg):
            groups.setdefault(cfg.dtype, []).append(i)
        else:
            solver = FlowSolver(cfg)
            solver.advance(int(steps))
            results[i] = dict(summary(solver), batched=False)
    for idx in groups.values():
        for b in range(0, len(idx), max_batch):
            part = idx[b:b + max_batch]
            ens = Ensemble2D([configs[i] for i in part])
            if tune and len(part) > 1:
                ens.tune()
            ens.advance(int(steps))
            for i, solver in zip(part, ens.solvers):
                results[i] = dict(summary(solver), batched=True)
    return results
