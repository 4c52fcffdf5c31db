"""CPU checks of the run loops (``aerolume_amd.runner``): chunk sizing, the blow-up and
dt-collapse checks of the reference runner (simulation/runner.py:444-600) on a scripted solver,
averaging start, cancellation, and the factorial order of sweeps (api.py:4877-4930)."""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

from aerolume_amd import runner as R


class ScriptedSolver:
    """Advances t by dt per step; dt and coefficients follow scripted functions of t."""

    def __init__(self, dt=0.01, dt_of_t=None, coeff_of_t=None):
        self.t, self.steps, self._dt = 0.0, 0, dt
        self.dt_of_t = dt_of_t or (lambda t: dt)
        self.coeff_of_t = coeff_of_t or (lambda t: np.array([1.0, 0.1, 0.0]))
        self.calls, self.averaging_at = [], None

    @property
    def dt(self):
        return self._dt

    def advance(self, n):
        self.calls.append(n)
        for _ in range(n):
            self._dt = self.dt_of_t(self.t)
            self.t += self._dt
            self.steps += 1

    def coefficients(self):
        return self.coeff_of_t(self.t)

    def start_averaging(self):
        self.averaging_at = self.t


def test_chunks_are_whole_blocks_after_a_short_first_chunk():
    s = ScriptedSolver(dt=0.01)
    out = R.run_chunked(s, 5.0, planned_dt=0.01, chunk_steps=137)
    assert s.calls[0] == R.FIRST_CHUNK_STEPS
    assert all(n % 10 == 0 for n in s.calls[:-1]) and max(s.calls) <= 137
    assert out["t"] >= 5.0 - 1e-9 and out["steps"] == sum(s.calls) and out["chunks"] == len(s.calls)
    assert not out["cancelled"]


def test_chunk_sizes_stop_at_the_target():
    assert R.chunk_sizes(0.0, 0.05, 0.01, 500, False, 10) == 5
    assert R.chunk_sizes(0.0, 10.0, 0.01, 500, True, 10) == R.FIRST_CHUNK_STEPS
    assert R.chunk_sizes(0.0, 10.0, 0.01, 137, False, 10) == 130
    assert R.chunk_sizes(0.0, 10.0, 0.01, 13, False, 10) == 13


def test_averaging_starts_at_t_avg_and_chunks_end_there():
    s = ScriptedSolver(dt=0.01)
    out = R.run_chunked(s, 3.0, planned_dt=0.01, t_avg=1.234, chunk_steps=500)
    assert s.averaging_at is not None and abs(s.averaging_at - 1.24) < 1e-9
    assert out["averaging_from"] == s.averaging_at


def test_nonfinite_or_huge_coefficients_stop_the_run():
    s = ScriptedSolver(coeff_of_t=lambda t: np.array([1.0, math.nan if t > 0.5 else 0.0, 0.0]))
    with pytest.raises(R.FlowDiverged, match="diverged"):
        R.run_chunked(s, 5.0, planned_dt=0.01)
    s = ScriptedSolver(coeff_of_t=lambda t: np.array([2e6 if t > 0.3 else 1.0]))
    with pytest.raises(R.FlowDiverged):
        R.run_chunked(s, 5.0, planned_dt=0.01)


def test_collapsed_time_step_stops_the_run():
    s = ScriptedSolver(dt=0.01, dt_of_t=lambda t: 0.01 if t < 0.4 else 5e-6)
    with pytest.raises(R.FlowDiverged, match="time step collapsed"):
        R.run_chunked(s, 5.0, planned_dt=0.01)
    # a step shrinking to 2e-3 of the plan is still healthy
    s = ScriptedSolver(dt=0.01, dt_of_t=lambda t: 0.01 if t < 0.4 else 2e-5)
    R.run_chunked(s, 0.41, planned_dt=0.01)


def test_cancel_stops_between_chunks():
    class Flag:
        def __init__(self):
            self.n = 0

        def is_set(self):
            self.n += 1
            return self.n > 2
    s = ScriptedSolver()
    out = R.run_chunked(s, 100.0, planned_dt=0.01, cancel=Flag())
    assert out["cancelled"] and out["chunks"] == 2


def test_sweep_points_follow_the_factorial_order():
    pts = R.sweep_points(alpha=[0, 4, 8], re=np.array([1e3, 1e4]), turbulence="none")
    want = [dict(turbulence="none", alpha=a, re=r) for a, r in itertools.product([0, 4, 8], [1e3, 1e4])]
    assert pts == want and isinstance(pts[0]["re"], float)
    with pytest.raises(ValueError):
        R.sweep_points(alpha=[], re=1.0)
    with pytest.raises(ValueError):
        R.sweep_points(re=1.0)


def test_batchable_rule():
    from tests import cases
    from aerolume_amd.ns2d.config import ThermalConfig
    import dataclasses
    cfg = cases.c1_cfg("float32")
    assert R._batchable(cfg)
    assert not R._batchable(dataclasses.replace(cfg, dye=True))
    assert not R._batchable(dataclasses.replace(cfg, thermal=ThermalConfig(prandtl=0.71, richardson=0.1)))
