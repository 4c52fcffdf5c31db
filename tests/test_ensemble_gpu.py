"""GPU parity of the batched 2D ensemble: every run of an ensemble of uneven grids equals the
same run stepped alone through aero2d_step, bit for bit (the batched launches share the
single-run kernel bodies), with and without HIP graphs."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("dtype,graphs,tune", [("float32", True, False), ("float64", False, False),
                                               ("float32", True, True)])
def test_ensemble_runs_equal_single_runs(dtype, graphs, tune):
    """With ``tune``: 9 steps, the autotuner, then 14 more; the tuner's timed launches leave the
    runs untouched and the single runs use the tile rows it chose (force sums round with them)."""
    _need_gpu()
    from aerolume_amd.ns2d.ensemble import Ensemble2D, sweep_configs
    from aerolume_amd.ns2d.solver import FlowSolver
    print(f"ensemble {dtype} graphs={graphs} tune={tune}", flush=True)
    cfgs = sweep_configs(n=5, seed=7, steps_cells=(48, 128), dtype=dtype)
    cfgs.append(dataclasses.replace(cfgs[0], nx=90, ny=45, h=8.0 / 90, perturb=True, re=500.0))   # odd ny, kick
    ens = Ensemble2D(cfgs, use_cuda_graphs=graphs)
    rows = 0
    if tune:
        ens.advance(9)
        choice = ens.tune(iters=2)
        print(f"ensemble tuner choice {choice}", flush=True)
        assert choice["tile_rows"] in (4, 8, 16) and choice["column_slots"] in (0, 1, 2, 4)
        assert choice["candidates"] == 48 and choice["row_threads"] in (256, 512) and choice["row_fill"] in (0, 1)
        rows = choice["tile_rows"]
        ens.advance(14)
    else:
        ens.advance(23)
    torch.cuda.synchronize()
    for c, se in zip(cfgs, ens.solvers):
        alone = FlowSolver(dataclasses.replace(c, use_cuda_graphs=False), tile_rows=rows)
        alone.advance(23)
        torch.cuda.synchronize()
        for f in ("u", "v", "p", "t", "dt"):
            assert torch.equal(getattr(se.state, f), getattr(alone.state, f)), (c.nx, f)
        he, ha = se.history(), alone.history()
        for k in he:
            if tune:
                # the 9 steps before the tuner summed forces over 8-row tiles, the single run
                # over `rows`: equal to round-off there, bit for bit after
                assert np.allclose(he[k][:9], ha[k][:9], rtol=1e-4, atol=1e-5), (c.nx, k)
                assert np.array_equal(he[k][9:], ha[k][9:]), (c.nx, k)
            else:
                assert np.array_equal(he[k], ha[k]), (c.nx, k)
        assert len(he["t"]) == 23
    assert ens.coefficients().shape == (len(cfgs), 3)
    assert all(np.isfinite(ens.coefficients()).all() for _ in [0])
