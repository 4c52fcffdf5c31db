"""The reference FlowSolver3D's kernel seam served by the HIP step.

The reference asks ``fused.load_kernels(scheme, device, dtype)`` for a kernel object when a
solver is built (ns3d/fused.py:226-251; a ``RuntimeError`` makes it fall back to its torch
core) and then calls ``fast_step(state, geom, ops, kern)`` or ``fast_step_phys(state, geom,
ops, kern, phys=..., theta=..., diag=...)`` once per step (ns3d/solver.py:583-595,
fused.py:371-502). The three functions here have the same names, arguments and results and run
one ``aero3d_step`` per call, so a maintainer swaps the import (the recipe is in the
integration notes next to this package).

What the step reads from the caller's objects is only the attributes the reference's own types
carry: ``Ops3D`` (grid, ``periodic_axes``, ``closed``, ``symmetry``, ``advection``,
``turbulence``, ``ghost``, ``cfl``, ``les_c``, ``ref_area`` / ``length_ref`` / ``span_ref`` /
``moment_ref``, the device scalars ``nu``, ``eta``, ``u_inlet``, ``extras``, ``dt_fixed``),
``Geometry3D`` (``window``, the masks, ``nbody``, ``empty``) and ``State3D`` (``u, v, w, p, t,
dt``). Heat transfer comes in as this package's :class:`~aerolume_amd.ns3d.thermal.ThermalOps3D`
(built once from the reference's config, ops, geometry and bodies, whose field names it
shares), passed as ``phys`` or as ``phys.thermal``; fans, ports and painted obstacles are not
carried and raise ``NotImplementedError``.
"""
from __future__ import annotations

import torch

from . import solver as _solver
from .thermal import ThermalOps3D


class HipKernels3D:
    """What :func:`load_kernels` returns: the scheme and device it was checked for. The
    plans themselves are cached on ``ops.extras`` by the functional steps."""

    def __init__(self, scheme: str, device: torch.device, dtype: torch.dtype):
        self.scheme, self.device, self.dtype = scheme, device, dtype


def load_kernels(scheme: str, device: torch.device, dtype: torch.dtype) -> HipKernels3D:
    """The kernel object for ``scheme`` on ``device`` / ``dtype``, or ``RuntimeError`` when the
    HIP step cannot run there (no GPU, library missing, unknown scheme or dtype), which the
    reference turns into its fallback (fused.py:226-251)."""
    device = torch.device(device)
    if device.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("the HIP step needs a GPU device")
    if scheme not in ("upwind3", "central"):
        raise RuntimeError(f"no HIP kernels for advection scheme {scheme!r}")
    if dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"no HIP kernels for dtype {dtype}")
    try:
        from .. import _lib
        _lib.load()
    except Exception as exc:  # noqa: BLE001 - any load failure means "not available"
        raise RuntimeError(f"HIP step unavailable ({type(exc).__name__}: {exc})") from exc
    return HipKernels3D(scheme, device, dtype)


def _as_state(state, like):
    """The new state in the caller's state type (the reference's NamedTuple or this one's)."""
    return type(like)(*state) if type(like) is not type(state) else state


def fast_step(state, geom, ops, kern: HipKernels3D, dt=None):
    """One step, ``(new state, coefficients (nbody, 6))`` (fused.py:371-375)."""
    new, coeffs = _solver.step3d_full(_solver.State3D(*state), geom, ops, dt)
    return _as_state(new, state), coeffs


def _thermal_of(phys):
    if phys is None:
        return None
    if isinstance(phys, ThermalOps3D):
        return phys
    for name in ("fans", "ports", "paint"):
        if getattr(phys, name, None):
            raise NotImplementedError(f"the HIP step does not carry {name}")
    th = getattr(phys, "thermal", None)
    if th is None or isinstance(th, ThermalOps3D):
        return th
    raise TypeError("heat transfer comes in as this package's ThermalOps3D (phys or phys.thermal)")


def fast_step_phys(state, geom, ops, kern: HipKernels3D, dt=None, *, phys=None, theta=None, diag: bool = False,
                   theta_kernels: bool = True) -> _solver.StepOut3D:
    """One step with heat transfer, ``StepOut3D(state, coefficients, theta)`` (fused.py:378-502):
    theta advances in the momentum's stages and its buoyancy acts on w; the heat rates and
    wall temperatures land in ``thermal.heat_rate`` / ``thermal.theta_wall``. ``diag`` and
    ``theta_kernels`` are accepted for the signature (diagnostics rows are the solver's)."""
    thermal = _thermal_of(phys)
    if thermal is not None and theta is None:
        raise ValueError("a heat-transfer step needs theta")
    out = _solver.step3d_phys(_solver.State3D(*state), geom, ops, dt, thermal=thermal, theta=theta)
    return _solver.StepOut3D(_as_state(out.state, state), out.coeffs, out.theta)
