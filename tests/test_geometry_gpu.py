"""The 2D mask builder on the MI355X: the HIP signed distance (``aero_polygon_sdf``) equals the
host restatement bit for bit, and the masks built on the GPU equal the reference's tensors
(``tests/golden/ns2d_masks.npz``) as the CPU-built ones do."""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np
import pytest
import torch

from aerolume_amd.ns2d import setup as su
from aerolume_amd.ns2d.ensemble import sweep_configs
from tests import cases

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_polygon_sdf_kernel_bitwise_equals_host(dtype):
    print(f"[geometry gpu] polygon sdf {dtype}", flush=True)
    g = torch.Generator().manual_seed(11)
    for n in (3, 7, 401, 2500):           # 2500: more than one LDS chunk of vertices
        th = torch.linspace(0, 2 * np.pi, n + 1, dtype=torch.float64)[:-1]
        r = 1.0 + 0.3 * torch.sin(5 * th) + 0.05 * torch.rand(n, generator=g, dtype=torch.float64)
        poly = torch.stack([0.3 + r * torch.cos(th), -0.1 + r * torch.sin(th)], 1).to(dtype)
        x = (torch.rand(3001, generator=g, dtype=torch.float64) * 4 - 2).to(dtype)
        y = (torch.rand(3001, generator=g, dtype=torch.float64) * 4 - 2).to(dtype)
        # points exactly on vertices and on a segment's line
        x[0], y[0] = poly[0, 0], poly[0, 1]
        x[1], y[1] = poly[1, 0], poly[1, 1]
        x[2], y[2] = 0.5 * (poly[0, 0] + poly[1, 0]), 0.5 * (poly[0, 1] + poly[1, 1])
        cpu = su.polygon_sdf(poly, x, y)
        gpu = su.polygon_sdf(poly, x.cuda(), y.cuda()).cpu()
        assert torch.equal(cpu, gpu), (n, float((cpu - gpu).abs().max()))
        assert (cpu < 0).any() and (cpu > 0).any()


def test_gpu_masks_match_reference():
    print("[geometry gpu] masks vs reference", flush=True)
    F = dict(np.load(GOLDEN / "ns2d_masks.npz"))
    for prefix, cfg, live in (("c1_", cases.c1_cfg("float64", device="cuda"), False),
                              ("c2_", cases.c2_cfg(dtype="float64", device="cuda"), True)):
        ops = su.make_operators(cfg)
        w = su.live_window(cfg.bodies, ops) if live else None
        g = su.build_geometry(cfg.bodies, ops, window=w)
        assert g.chi_u.is_cuda and tuple(g.window) == tuple(F[prefix + "window"])
        for k in ("chi_u", "chi_v", "chi_c", "a_u", "a_v", "k_u", "k_v"):
            assert np.abs(getattr(g, k).cpu().numpy() - F[prefix + k]).max() < 1e-14, (prefix, k)
        if prefix == "c2_":
            for k in ("chi_t", "a_t", "w_t", "vol_u", "vol_v", "ins_u", "ins_v"):
                assert np.abs(getattr(g.thermal, k).cpu().numpy() - F[prefix + "th_" + k]).max() < 1e-14, k


def test_sweep_masks_gpu_equal_cpu_and_timing():
    """The 64 airfoil masks of the batched sweep built on the GPU equal the CPU build (float32:
    distances bit for bit, masks to the sine's last-place rounding); prints both build times."""
    print("[geometry gpu] sweep masks", flush=True)
    cfgs = sweep_configs(64, device="cuda")
    ops_g = [su.make_operators(c) for c in cfgs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    geo_g = [su.build_geometry(c.bodies, o, window=su.live_window(c.bodies, o)) for c, o in zip(cfgs, ops_g)]
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    ops_c = [su.make_operators(c, device="cpu") for c in cfgs]
    geo_c = [su.build_geometry(c.bodies, o, window=su.live_window(c.bodies, o)) for c, o in zip(cfgs, ops_c)]
    t_cpu = time.perf_counter() - t0
    print(f"[geometry gpu] 64 sweep geometries: gpu {t_gpu * 1e3:.1f} ms, cpu {t_cpu * 1e3:.1f} ms "
          f"({torch.get_num_threads()} threads)", flush=True)
    for a, b in zip(geo_g, geo_c):
        assert a.window == b.window
        for k in ("chi_u", "chi_v", "chi_c", "k_u", "k_v"):
            assert float((getattr(a, k).cpu() - getattr(b, k)).abs().max()) <= 4e-7, k
