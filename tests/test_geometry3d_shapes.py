"""The shape descriptors the 3D analytic bodies carry for the HIP fraction kernel describe the
same bodies as their signed-distance closures: the distance restated from the descriptor's
seven parameters (numpy, the kernel's formulas) equals the closure's on random points, and the
descriptor packs into the C struct the library takes."""
from __future__ import annotations

import numpy as np
import torch


def _from_shape(kind, p, x, y, z):
    if kind == "rounded_box":
        q = [np.abs(x - p[0]) - p[3], np.abs(y - p[1]) - p[4], np.abs(z - p[2]) - p[5]]
        out = np.sqrt(sum(np.maximum(c, 0.0) ** 2 for c in q))
        return out + np.minimum(np.maximum(np.maximum(q[0], q[1]), q[2]), 0.0) - p[6]
    dr = np.sqrt((x - p[0]) ** 2 + (z - p[2]) ** 2) - p[3]
    da = np.abs(y - p[1]) - p[4]
    return np.sqrt(np.maximum(dr, 0.0) ** 2 + np.maximum(da, 0.0) ** 2) + np.minimum(np.maximum(dr, da), 0.0)


def test_descriptors_match_closures():
    from aerolume_amd.ns3d.setup import car_bodies, cylinder_y_sdf, rounded_box_sdf
    rng = np.random.default_rng(7)
    pts = rng.uniform(-3.0, 3.0, size=(3, 4000))
    sdfs = [b.sdf for b in car_bodies(center=(0.3, -0.2, 0.1))]
    sdfs += [rounded_box_sdf((0.5, 1.0, -0.5), (2.0, 0.7, 1.3), 0.3), cylinder_y_sdf((1.0, -1.0, 0.5), 0.8, 1.6)]
    for sdf in sdfs:
        kind, params = sdf.shape
        assert len(params) == 7
        want = sdf(*(torch.as_tensor(c) for c in pts)).numpy()
        np.testing.assert_allclose(_from_shape(kind, params, *pts), want, rtol=0, atol=1e-14)


def test_descriptor_packs_into_struct():
    from aerolume_amd._lib3d import SHAPE_KINDS, Shape3D
    from aerolume_amd.ns3d.setup import cylinder_y_sdf
    kind, params = cylinder_y_sdf((1.0, 2.0, 3.0), 0.5, 0.8).shape
    s = Shape3D()
    s.kind = SHAPE_KINDS[kind]
    for q, v in enumerate(params):
        s.p[q] = v
    assert s.kind == 1 and list(s.p) == [1.0, 2.0, 3.0, 0.5, 0.4, 0.0, 0.0]
