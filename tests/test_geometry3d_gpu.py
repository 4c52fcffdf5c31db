"""GPU 3D masks: the analytic-shape fraction kernel (``aero3d_analytic_fractions``) behind
``build_geometry3d`` on a GPU device gives the host builder's masks, rotation fields and force
weights (float64, within 1e-15 of each array's magnitude: the two sides evaluate the same formulas,
and only the sine's last bit may differ), for the car case, overlapping bodies, more
than one launch's worth of bodies, and windows touching the domain boundary. A body without a
known shape keeps the host path on the GPU too."""
from __future__ import annotations

import dataclasses

import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _both(cfg):
    """Geometry built for the GPU device and for the host, the GPU one moved to the host."""
    from aerolume_amd.ns3d import setup as su
    gpu_ops = su.make_operators3d(dataclasses.replace(cfg, device="cuda"))
    cpu_ops = su.make_operators3d(dataclasses.replace(cfg, device="cpu"))
    g = su.build_geometry3d(cfg.bodies, gpu_ops)
    c = su.build_geometry3d(cfg.bodies, cpu_ops)
    return g, c


def _compare(g, c, tol=1e-15):
    assert tuple(g.window) == tuple(c.window) and g.nbody == c.nbody
    for name in ("chi_u", "chi_v", "chi_w", "chi_c", "a_u", "a_v", "a_w", "k_u", "k_v", "k_w"):
        a, b = getattr(g, name), getattr(c, name)
        assert a.is_cuda and a.dtype == b.dtype and a.shape == b.shape, name
        scale = max(1.0, float(b.abs().max()))
        err = float((a.cpu() - b).abs().max())
        assert err <= tol * scale, (name, err)


def test_car_masks_equal_host_builder():
    _need_gpu()
    print("[geometry3d gpu] car", flush=True)
    g, c = _both(cases.c3_cfg("float64", device="cuda", nx=96, ny=48, nz=48))
    _compare(g, c)
    assert float(g.chi_c.max()) == 1.0 and float(g.a_u.abs().max()) > 0.0


@pytest.mark.parametrize("name", ["laminar", "wale_moving_py", "vreman_symmetry"])
def test_small_cases_equal_host_builder(name):
    _need_gpu()
    print(f"[geometry3d gpu] {name}", flush=True)
    g, c = _both(cases.cfg3d(name, "float64"))
    _compare(g, c)


def test_many_overlapping_bodies_equal_host_builder():
    """18 overlapping shapes (two launches of at most 16), shares where they overlap."""
    _need_gpu()
    from aerolume_amd.ns3d.setup import AnalyticBody, cylinder_y_sdf, rounded_box_sdf
    print("[geometry3d gpu] 18 bodies", flush=True)
    bodies = []
    for q in range(18):
        x, z = -0.3 + 0.07 * q, 0.6 + 0.02 * (q % 5)
        if q % 3 == 2:
            c = (x, 0.05, z)
            bodies.append(AnalyticBody(cylinder_y_sdf(c, 0.12, 0.3), ((x - 0.12, -0.1, z - 0.12),
                                                                       (x + 0.12, 0.2, z + 0.12)),
                                       omega=1.5 + q, center=c, axis=(0.0, 1.0, 0.0)))
        else:
            bodies.append(AnalyticBody(rounded_box_sdf((x, 0.0, z), (0.2, 0.3, 0.25), 0.05),
                                       ((x - 0.1, -0.15, z - 0.125), (x + 0.1, 0.15, z + 0.125))))
    cfg = dataclasses.replace(cases.cfg3d("laminar", "float64"), bodies=bodies)
    g, c = _both(cfg)
    _compare(g, c)
    assert g.k_u.shape[0] == 18


def test_unknown_shape_keeps_host_path():
    _need_gpu()
    from aerolume_amd.ns3d.setup import AnalyticBody
    print("[geometry3d gpu] custom sdf", flush=True)

    def ball(x, y, z):
        return torch.sqrt((x - 0.2) ** 2 + y ** 2 + (z - 0.9) ** 2) - 0.25
    cfg = dataclasses.replace(cases.cfg3d("laminar", "float64"),
                              bodies=[AnalyticBody(ball, ((-0.05, -0.25, 0.65), (0.45, 0.25, 1.15)))])
    g, c = _both(cfg)
    for name in ("chi_u", "chi_v", "chi_w", "chi_c", "k_u"):
        assert torch.equal(getattr(g, name).cpu(), getattr(c, name)), name
