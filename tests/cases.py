"""Shared case builders of the tests (shapes and configs of the benchmark cases)."""
from __future__ import annotations

import math

import numpy as np


def circle(n, d, cx, cy):
    th = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    return np.c_[cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)]


def naca4(code="0012", n=400, chord=1.0):
    """Symmetric NACA 4-digit section, closed trailing edge, CCW (TE, upper, LE, lower)."""
    t = int(code[2:]) / 100
    beta = np.linspace(0.0, np.pi, n // 2 + 1)
    x = 0.5 * (1 - np.cos(beta))
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 - 0.1036 * x**4)
    return np.vstack([np.c_[x, yt][::-1], np.c_[x, -yt][1:-1]]) * chord


def rotate(pts, deg, about=(0.0, 0.0)):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    p = np.asarray(pts, float) - about
    return np.c_[c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1]] + about


def trace_cfg(dtype="float64", device="cpu"):
    from aerolume_amd.ns2d.config import ImmersedBody, SolverConfig
    return SolverConfig(nx=64, ny=32, h=1.0 / 8.0, x0=0.0, y0=0.0, re=100.0,
                        bodies=[ImmersedBody(circle(64, 0.75, 2.0, 2.0))], dtype=dtype, device=device,
                        dye=False, perturb=False)


def c1_cfg(dtype="float32", device="cpu", nx=256, ny=128):
    """2D plain case: 256x128, domain 8x4, disc D=0.5 at (2, 2), Re 100, upwind3, no LES."""
    from aerolume_amd.ns2d.config import ImmersedBody, SolverConfig
    return SolverConfig(nx=nx, ny=ny, h=8.0 / nx, x0=0.0, y0=0.0, re=100.0,
                        bodies=[ImmersedBody(circle(256, 0.5, 2.0, 2.0))], advection="upwind3",
                        turbulence="none", cfl=0.5, dtype=dtype, device=device, dye=False, perturb=False)


def c2_cfg(nx=1024, ny=512, dtype="float32", device="cpu", turbulence="smagorinsky"):
    """Heated airfoil: NACA 0012 at 8 deg held at theta = 1, Re 1e5, Smagorinsky, Pr 0.71,
    Ri 0.1, one line smoke emitter; domain [-2, 6] x [-2, 2]."""
    from aerolume_amd.ns2d.config import EmitterSpec, ImmersedBody, SolverConfig, ThermalConfig
    foil = rotate(naca4("0012", 400), -8.0, about=(0.25, 0.0))
    return SolverConfig(nx=nx, ny=ny, h=8.0 / nx, x0=-2.0, y0=-2.0, re=1e5,
                        bodies=[ImmersedBody(foil, thermal="temperature", thermal_value=1.0)],
                        advection="upwind3", turbulence=turbulence, cfl=0.5, dtype=dtype, device=device,
                        dye=True, dye_inlet=False, perturb=False,
                        thermal=ThermalConfig(prandtl=0.71, richardson=0.1),
                        emitters=[EmitterSpec(kind="line", position=(-1.0, -0.2), end=(-1.0, 0.2), channel=0,
                                              rate=1.0)])
