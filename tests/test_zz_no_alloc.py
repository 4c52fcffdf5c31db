"""The allocation-free Poisson solve (``aero2d_poisson_solve_static``: scratch in a static
device array, no plan) against the plan path (``aero2d_poisson_solve``) through the C ABI:
two calls in a row on one GPU, the second on a smaller grid, in float32 and float64.

Bar: the two paths run the same transform kernels; only the tables differ in where they are
computed (device ``cos``/``sin`` against the host's), so they agree to within a few ulps of
the result: 1e-12 relative in float64, 2e-6 relative in float32."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _plan_solve(f, h, periodic):
    from aerolume_amd.ns2d import solver as S
    from aerolume_amd.ns2d import setup as su
    from aerolume_amd.ns2d.config import SolverConfig
    ny, nx = f.shape
    dtype = "float64" if f.dtype == torch.float64 else "float32"
    cfg = SolverConfig(nx=nx, ny=ny, h=h, x0=0, y0=0, re=100, dtype=dtype, device="cpu", dye=False, perturb=False,
                       periodic=periodic)
    ops = S.make_operators(cfg, device=DEV)
    geom = su.build_geometry(cfg.bodies, ops)
    psi = torch.empty_like(f)
    S.get_plan(ops, geom).poisson(f, psi)
    return psi


def _static_solve(f, h, periodic, psi=None):
    from aerolume_amd import _lib
    lib = _lib.load()
    ny, nx = f.shape
    psi = torch.empty_like(f) if psi is None else psi
    st = _lib.stream_handle(DEV)
    rc = lib.aero2d_poisson_solve_static(ctypes.c_void_p(st), 1 if f.dtype == torch.float64 else 0, nx, ny, h,
                                         1 if periodic else 0, f.data_ptr(), psi.data_ptr())
    _lib.check(rc, "aero2d_poisson_solve_static")
    return psi


def _rhs(ny, nx, dtype, seed):
    f = torch.as_tensor(np.random.default_rng(seed).standard_normal((ny, nx)), dtype=dtype, device=DEV)
    return f - f.mean()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_static_solve_equals_plan_solve_twice(dtype):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    # first call at the 1024 x 512 airfoil grid, second on a smaller periodic grid: the second
    # call must not see anything the first left in the shared array
    for (nx, ny, h, periodic, seed) in ((1024, 512, 8.0 / 1024, False, 11), (96, 60, 0.1, True, 12)):
        print(f"static poisson {nx}x{ny} periodic={periodic} {dtype}", flush=True)
        f = _rhs(ny, nx, dtype, seed)
        ref = _plan_solve(f, h, periodic)
        got = _static_solve(f, h, periodic)
        torch.cuda.synchronize()
        scale = float(ref.abs().max())
        err = float((got - ref).abs().max())
        assert err <= tol * scale, (nx, ny, err, scale)


def test_static_solve_long_rows_and_in_place():
    """Rows of 1200 float64 values (more than 64 KB of LDS per block in the row passes), and
    psi written over f (the entry point allows aliasing)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("static poisson 1200x24 periodic float64 in place", flush=True)
    f = _rhs(24, 1200, torch.float64, 13)
    ref = _plan_solve(f, 0.05, True)
    got = _static_solve(f, 0.05, True, psi=f)
    torch.cuda.synchronize()
    assert got.data_ptr() == f.data_ptr()
    assert float((got - ref).abs().max()) <= 1e-12 * float(ref.abs().max())


def test_static_solve_refuses_grids_beyond_its_array():
    """Sizes beyond the static array's limit are refused on the host, before any launch."""
    from aerolume_amd import _lib
    lib = _lib.load()
    for nx, ny in ((2048, 1024), (8192, 4)):
        rc = lib.aero2d_poisson_solve_static(None, 0, nx, ny, 0.1, 0, ctypes.c_void_p(16), ctypes.c_void_p(16))
        assert rc == -2
        assert b"static scratch" in lib.aero_last_error()
    assert lib.aero2d_poisson_solve_static(None, 0, 64, 32, 0.1, 0, None, None) == -1
