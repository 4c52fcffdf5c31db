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


@pytest.mark.parametrize("dtype,graphs", [("float32", True), ("float64", False)])
def test_ensemble_runs_equal_single_runs(dtype, graphs):
    _need_gpu()
    from aerolume_amd.ns2d.ensemble import Ensemble2D, sweep_configs
    from aerolume_amd.ns2d.solver import FlowSolver
    print(f"ensemble {dtype} graphs={graphs}", flush=True)
    cfgs = sweep_configs(n=5, seed=7, steps_cells=(48, 128), dtype=dtype)
    cfgs.append(dataclasses.replace(cfgs[0], nx=90, ny=45, h=8.0 / 90, perturb=True, re=500.0))   # odd ny, kick
    ens = Ensemble2D(cfgs, use_cuda_graphs=graphs)
    ens.advance(23)
    torch.cuda.synchronize()
    for c, se in zip(cfgs, ens.solvers):
        alone = FlowSolver(dataclasses.replace(c, use_cuda_graphs=False))
        alone.advance(23)
        torch.cuda.synchronize()
        for f in ("u", "v", "p", "t", "dt"):
            assert torch.equal(getattr(se.state, f), getattr(alone.state, f)), (c.nx, f)
        he, ha = se.history(), alone.history()
        for k in he:
            assert np.array_equal(he[k], ha[k]), (c.nx, k)
        assert len(he["t"]) == 23
    assert ens.coefficients().shape == (len(cfgs), 3)
    assert all(np.isfinite(ens.coefficients()).all() for _ in [0])
