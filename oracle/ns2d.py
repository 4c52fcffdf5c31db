"""CPU oracle of the 2D Navier-Stokes step (TEST INFRASTRUCTURE, never the product).

A restatement, in plain torch on the CPU (float32 or float64), of the reference's 2D
time step ``step_full`` (``solvers/ns2d/core.py:815-969``) and every helper it calls
for the features the MI355X step carries: tunnel / periodic / closed domains, upwind3 and
central advection, Smagorinsky / WALE / Vreman eddy viscosity, window penalisation with
forces, the exact DCT / FFT projection, temperature transport with Boussinesq buoyancy
and heat outputs, and smoke (inlet comb or emitters).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use this
module, as the checker. Each function cites the reference lines it follows. It takes the
set-up objects (``Ops``, ``Geometry``, ``Physics``) as plain data.

Pinning (what checks each function besides retyping it):
* ``PoissonSolver`` - discrete Laplacian residual and zero mean on odd / even sizes, both
  BC kinds (closed form), scipy ``dctn`` (textbook), dense least squares on tiny grids.
* ``project`` - divergence of white noise to round-off, boundary-normal faces untouched.
* ``step_full`` and helpers - the reference's own trajectories (golden fixtures written by
  ``tests/golden/make_ns2d_fixtures.py``), Taylor-Green decay, energy conservation of the
  central scheme, quadratic exactness of the upwind3 reconstruction.
* LES models - closed forms in float64 numpy, zero in pure shear for WALE / Vreman.
* thermal, smoke - the heated-airfoil fixture, conduction / maximum-principle invariants.
* ``init_state``, ``advect_dye`` - the reference FlowSolver's default path (Lamb-Oseen start,
  inlet smoke comb; ``tests/golden/ns2d_default_*.npz`` written by
  ``tests/golden/make_ns2d_default_fixtures.py``), a planted zero-inflow comb, and the comb's
  invariants (conservation, no new extrema, uniform stays uniform).
* ``pid_dt`` - the reference's PID trajectory (``tests/golden/ns2d_pid_diag.npz``, written by
  ``tests/golden/make_ns2d_diag_fixtures.py``) and the controller's closed forms (first step,
  fixed point at the target, growth cap, clip at the limit).
* ``diagnostics_row`` - the reference FlowSolver's rows at steps 10 and 20 (same fixture), the
  residual column's closed forms (exact potential, half of it, an incompatible right-hand
  side), a planted change norm without its h factor.
"""
from __future__ import annotations

import math
from typing import NamedTuple

import torch

SSP_RK3 = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))     # core.py:47
FORCE_WEIGHTS = (1.0 / 6.0, 2.0 / 3.0, 1.0)                      # core.py:49
SQRT_EPS = 1e-30                                                 # core.py:441
DYE_FLOOR = 1e-6                                                 # scalar.py:31


class State(NamedTuple):
    """Evolving fields (core.py:112-129)."""
    u: torch.Tensor
    v: torch.Tensor
    p: torch.Tensor
    dye: torch.Tensor | None
    t: torch.Tensor
    dt: torch.Tensor
    theta: torch.Tensor | None = None
    dye_extra: torch.Tensor | None = None
    ctrl: torch.Tensor | None = None


# ---------------------------------------------------------------------------
# Poisson (poisson.py:45-144): Makhoul DCT-II via a real FFT, natural-order coefficients
# ---------------------------------------------------------------------------

def neumann_eigenvalues(n: int, h: float) -> torch.Tensor:
    """(2 cos(pi k / n) - 2) / h^2 in float64 (poisson.py:90-92)."""
    return torch.tensor([(2.0 * math.cos(math.pi * k / n) - 2.0) / h**2 for k in range(n)], dtype=torch.float64)


def _reorder(n: int) -> tuple[torch.Tensor, torch.Tensor]:
    perm = list(range(0, n, 2)) + list(range(1, n, 2))[::-1]        # poisson.py:57
    inv = [0] * n
    for pos, src in enumerate(perm):
        inv[src] = pos
    return torch.tensor(perm), torch.tensor(inv)


def dct2(x: torch.Tensor, dim: int) -> torch.Tensor:
    """X_k = sum_n x_n cos(pi k (2n+1) / 2N) along ``dim`` (Makhoul: reorder, FFT, twiddle)."""
    n = x.shape[dim]
    perm, _ = _reorder(n)
    v = x.index_select(dim, perm)
    V = torch.fft.fft(v.to(torch.complex128 if x.dtype == torch.float64 else torch.complex64), dim=dim)
    k = torch.arange(n, dtype=torch.float64)
    tw = torch.polar(torch.ones_like(k), -math.pi * k / (2 * n)).to(V.dtype)
    shape = [1] * x.dim()
    shape[dim] = n
    return (V * tw.reshape(shape)).real.to(x.dtype)


def idct2(X: torch.Tensor, dim: int) -> torch.Tensor:
    """Exact inverse of :func:`dct2`: V_k = e^{i pi k / 2N} (X_k - i X_{N-k}), IFFT, un-reorder."""
    n = X.shape[dim]
    _, iperm = _reorder(n)
    cd = torch.complex128 if X.dtype == torch.float64 else torch.complex64
    Xr = torch.cat([torch.zeros_like(X.narrow(dim, 0, 1)), X.narrow(dim, 1, n - 1).flip(dim)], dim=dim)
    k = torch.arange(n, dtype=torch.float64)
    tw = torch.polar(torch.ones_like(k), math.pi * k / (2 * n)).to(cd)
    shape = [1] * X.dim()
    shape[dim] = n
    V = torch.complex(X, -Xr) * tw.reshape(shape)
    v = torch.fft.ifft(V, dim=dim).real.to(X.dtype)
    return v.index_select(dim, iperm)


class PoissonSolver:
    """Exact solve of the 5-point cell-centred Laplacian: homogeneous Neumann (DCT-II along
    both axes) or doubly periodic (real FFT); zero mode set to zero (poisson.py:95-134).
    The eigenvalue inverse is formed in float64 and cast (poisson.py:119-126)."""

    def __init__(self, nx: int, ny: int, h: float, *, periodic: bool, dtype: torch.dtype):
        self.nx, self.ny, self.h, self.periodic = nx, ny, h, periodic
        if periodic:
            kx = torch.arange(nx // 2 + 1, dtype=torch.float64)
            ky = torch.arange(ny, dtype=torch.float64)
            lx = (2 * torch.cos(2 * math.pi * kx / nx) - 2) / h**2
            ly = (2 * torch.cos(2 * math.pi * ky / ny) - 2) / h**2
        else:
            lx = neumann_eigenvalues(nx, h)
            ly = neumann_eigenvalues(ny, h)
        lam = ly[:, None] + lx[None, :]
        inv = torch.where(lam != 0, 1.0 / torch.where(lam != 0, lam, 1.0), 0.0)
        inv[0, 0] = 0.0
        self.inv_lam = inv.to(dtype)

    def solve(self, f: torch.Tensor) -> torch.Tensor:
        if self.periodic:
            fh = torch.fft.rfft2(f) * self.inv_lam
            return torch.fft.irfft2(fh, s=(self.ny, self.nx))
        fh = dct2(dct2(f, 1), 0) * self.inv_lam
        return idct2(idct2(fh, 0), 1)

    def laplacian(self, psi: torch.Tensor) -> torch.Tensor:
        """The discrete operator itself (poisson.py:136-144)."""
        if self.periodic:
            pe = torch.cat([psi[:, -1:], psi, psi[:, :1]], dim=1)
            pe = torch.cat([pe[-1:], pe, pe[:1]], dim=0)
        else:
            pe = torch.cat([psi[:, :1], psi, psi[:, -1:]], dim=1)
            pe = torch.cat([pe[:1], pe, pe[-1:]], dim=0)
        return (pe[1:-1, 2:] + pe[1:-1, :-2] + pe[2:, 1:-1] + pe[:-2, 1:-1] - 4 * pe[1:-1, 1:-1]) / self.h**2


_POISSON: dict = {}


def poisson(ops) -> PoissonSolver:
    key = (ops.nx, ops.ny, ops.h, ops.periodic, ops.dtype)
    if key not in _POISSON:
        _POISSON[key] = PoissonSolver(ops.nx, ops.ny, ops.h, periodic=ops.periodic, dtype=ops.dtype)
    return _POISSON[key]


# ---------------------------------------------------------------------------
# ghosts (core.py:278-357)
# ---------------------------------------------------------------------------

def _wall_ghost(rows, kind, speed):
    if kind == "slip":
        return rows
    if kind == "noslip":
        return -rows
    return 2.0 * speed - rows


def pad_u(u, ops):
    """u with two ghost layers, (ny+4, nx+5) (core.py:287-305)."""
    if ops.periodic:
        up = u[:, :-1]
        up = torch.cat([up[:, -2:], up, up[:, :3]], dim=1)
        return torch.cat([up[-2:], up, up[:2]], dim=0)
    if ops.closed:
        ux = torch.cat([-u[:, 1:3].flip(1), u, -u[:, -3:-1].flip(1)], dim=1)
    else:
        ux = torch.cat([u[:, :1].expand(-1, 2), u, u[:, -1:].expand(-1, 2)], dim=1)
    bot = _wall_ghost(ux[:2].flip(0), ops.bc_bottom, ops.ground_speed)
    top = _wall_ghost(ux[-2:].flip(0), ops.bc_top, 0.0)
    return torch.cat([bot, ux, top], dim=0)


def pad_v(v, ops):
    """v with two ghost layers, (ny+5, nx+4) (core.py:308-325)."""
    if ops.periodic:
        vp = v[:-1]
        vp = torch.cat([vp[-2:], vp, vp[:3]], dim=0)
        return torch.cat([vp[:, -2:], vp, vp[:, :2]], dim=1)
    vy = torch.cat([-v[1:3].flip(0), v, -v[-3:-1].flip(0)], dim=0)
    if ops.closed:
        sl = -1.0 if ops.bc_left == "noslip" else 1.0
        sr = -1.0 if ops.bc_right == "noslip" else 1.0
        return torch.cat([sl * vy[:, :2].flip(1), vy, sr * vy[:, -2:].flip(1)], dim=1)
    return torch.cat([-vy[:, :2].flip(1), vy, vy[:, -1:].expand(-1, 2)], dim=1)


def apply_boundary_values(u, v, ops):
    """Dirichlet inlet u, wall v, outlet balance (core.py:328-334)."""
    if ops.periodic:
        return torch.cat([u[:, :-1], u[:, :1]], dim=1), torch.cat([v[:-1], v[:1]], dim=0)
    u = inlet_outlet(u, ops)
    zero = torch.zeros_like(v[:1])
    return u, torch.cat([zero, v[1:-1], zero], dim=0)


def inlet_outlet(u, ops):
    """Inlet column from the profile; outlet shifted so outflow = inflow (core.py:337-357)."""
    if ops.closed:
        zc = torch.zeros_like(u[:, :1])
        return torch.cat([zc, u[:, 1:-1], zc], dim=1)
    corr = (ops.u_inlet.sum() - u[:, -1].sum()) / ops.ny
    return torch.cat([ops.u_inlet, u[:, 1:-1], u[:, -1:] + corr], dim=1)


# ---------------------------------------------------------------------------
# tendencies (core.py:364-503, 571-614)
# ---------------------------------------------------------------------------

def face_flux(vsum, a, b, c, d, upwind):
    """Scaled face flux (core.py:364-388): central vsum (b+c); upwind3
    vsum [(b+c) - (a+d)/7] + 3/7 |vsum| [(b-c) - (a-d)/3]."""
    s = b + c
    vsum = s if vsum is None else vsum
    if not upwind:
        return vsum * s
    cen = torch.add(s, a + d, alpha=-1.0 / 7.0)
    dis = torch.add(b - c, a - d, alpha=-1.0 / 3.0)
    return torch.addcmul(vsum * cen, vsum.abs(), dis, value=3.0 / 7.0)


def advection(U, V, ops):
    """-div(u u) at u- and v-faces (core.py:391-408)."""
    ny, nx = ops.ny, ops.nx
    up = ops.advection == "upwind3"
    k = (7.0 / 24.0 if up else 0.25) / ops.h
    fx = face_flux(None, U[2:-2, 0:nx + 2], U[2:-2, 1:nx + 3], U[2:-2, 2:nx + 4], U[2:-2, 3:nx + 5], up)
    vs = V[2:ny + 3, 1:nx + 2] + V[2:ny + 3, 2:nx + 3]
    gy = face_flux(vs, U[0:ny + 1, 2:-2], U[1:ny + 2, 2:-2], U[2:ny + 3, 2:-2], U[3:ny + 4, 2:-2], up)
    au = (fx[:, :-1] - fx[:, 1:]).add_(gy[:-1]).sub_(gy[1:]).mul_(k)
    us = U[1:ny + 2, 2:-2] + U[2:ny + 3, 2:-2]
    fx = face_flux(us, V[2:-2, 0:nx + 1], V[2:-2, 1:nx + 2], V[2:-2, 2:nx + 3], V[2:-2, 3:nx + 4], up)
    gy = face_flux(None, V[0:ny + 2, 2:-2], V[1:ny + 3, 2:-2], V[2:ny + 4, 2:-2], V[3:ny + 5, 2:-2], up)
    av = (fx[:, :-1] - fx[:, 1:]).add_(gy[:-1]).sub_(gy[1:]).mul_(k)
    return au, av


def add_diffusion(acc, P, nu_h2):
    """acc + nu/h^2 * 5-point ring (core.py:411-418)."""
    ring = (P[2:-2, 3:-1] + P[2:-2, 1:-3]).add_(P[3:-1, 2:-2]).add_(P[1:-3, 2:-2])
    return torch.addcmul(acc, ring.add_(P[2:-2, 2:-2], alpha=-4.0), nu_h2)


def gradients(U, V, ops):
    """Gradient tensor at cells -1..ny x -1..nx and corner shear (core.py:421-436)."""
    ny, nx, inv_h = ops.ny, ops.nx, 1.0 / ops.h
    g11 = (U[1:ny + 3, 2:nx + 4] - U[1:ny + 3, 1:nx + 3]) * inv_h
    g22 = (V[2:ny + 4, 1:nx + 3] - V[1:ny + 3, 1:nx + 3]) * inv_h
    dudy = (U[1:ny + 4, 1:nx + 4] - U[0:ny + 3, 1:nx + 4]) * inv_h
    dvdx = (V[1:ny + 4, 1:nx + 4] - V[1:ny + 4, 0:nx + 3]) * inv_h

    def avg(q):
        return (q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]) * 0.25
    return g11, avg(dudy), avg(dvdx), g22, dudy + dvdx


def eddy_viscosity(g11, g12, g21, g22, ops):
    """Smagorinsky / WALE / Vreman nu_t (core.py:444-478)."""
    delta2 = ops.h * ops.h
    c = ops.les_c
    s12 = 0.5 * (g12 + g21)
    ss = g11 * g11 + g22 * g22 + 2.0 * s12 * s12
    if ops.turbulence == "smagorinsky":
        return (c * c * delta2) * torch.sqrt(2.0 * ss + SQRT_EPS)
    if ops.turbulence == "wale":
        q11 = g11 * g11 + g12 * g21
        q22 = g21 * g12 + g22 * g22
        q12 = 0.5 * (g11 * g12 + g12 * g22 + g21 * g11 + g22 * g21)
        tr3 = (q11 + q22) * (1.0 / 3.0)
        sd = (q11 - tr3) ** 2 + (q22 - tr3) ** 2 + 2.0 * q12 * q12 + tr3 * tr3
        num = sd.clamp_min(0.0) ** 1.5
        den = ss.clamp_min(0.0) ** 2.5 + sd.clamp_min(0.0) ** 1.25 + 1e-30
        return (c * c * delta2) * num / den
    if ops.turbulence == "vreman":
        b11 = g11 * g11 + g12 * g12
        b22 = g21 * g21 + g22 * g22
        b12 = g11 * g21 + g12 * g22
        bb = (b11 * b22 - b12 * b12).clamp_min(0.0)
        aa = g11 * g11 + g12 * g12 + g21 * g21 + g22 * g22
        return (c * delta2) * torch.sqrt(bb / (aa + 1e-30) + SQRT_EPS)
    raise ValueError(ops.turbulence)


def sgs_stress_divergence(U, V, geom, ops):
    """div(2 nu_t S) at faces, max nu_t and the nu_t field (core.py:481-503)."""
    ny, nx, inv_h = ops.ny, ops.nx, 1.0 / ops.h
    g11, g12, g21, g22, shear_k = gradients(U, V, ops)
    nut = eddy_viscosity(g11, g12, g21, g22, ops)
    if not geom.empty:
        j0, j1, i0, i1 = geom.window
        win = nut[j0 + 1:j1 + 1, i0 + 1:i1 + 1] * (1.0 - geom.chi_c)
        nut = nut.clone()
        nut[j0 + 1:j1 + 1, i0 + 1:i1 + 1] = win
    nut_k = (nut[:-1, :-1] + nut[1:, :-1] + nut[:-1, 1:] + nut[1:, 1:]) * 0.25
    txx = 2.0 * nut[1:ny + 1] * g11[1:ny + 1]
    tyy = 2.0 * nut[:, 1:nx + 1] * g22[:, 1:nx + 1]
    txy = nut_k * shear_k[1:ny + 2, 1:nx + 2]
    fu = (txx[:, 1:] - txx[:, :-1] + txy[1:] - txy[:-1]) * inv_h
    fv = (txy[:, 1:] - txy[:, :-1] + tyy[1:] - tyy[:-1]) * inv_h
    return fu, fv, nut.amax(), nut


WW_A = 8.3          # Werner-Wengle power law u+ = A (y+)^B (core.py:509-513)
WW_B = 1.0 / 7.0


def wall_shear_excess(us, y, nu):
    """The Werner-Wengle wall stress above the resolved no-slip one nu |us| / y, signed like
    ``us`` (core.py:516-533): tau = (|us| (nu / y)^B / A)^(2 / (1 + B)), clamped at zero below
    the crossover y+ = A^(1/(1-B))."""
    a = us.abs()
    tau_lam = a * (nu / y)
    tau_ww = (a * (nu / y) ** WW_B * (1.0 / WW_A)) ** (2.0 / (1.0 + WW_B))
    return torch.sign(us) * (tau_ww - tau_lam).clamp_min(0.0)


def wall_model_force(u, v, ops):
    """Force per unit volume (-excess / h) on the u faces next to no-slip (moving: relative to
    the wall speed) bottom / top walls, and with a closed domain on the v faces next to
    no-slip side walls (core.py:536-568); v part None when no side wall is modelled."""
    y = 0.5 * ops.h
    inv_h = 1.0 / ops.h
    fu = torch.zeros_like(u)
    if ops.bc_bottom in ("noslip", "moving"):
        us = u[:1] - ops.ground_speed if ops.bc_bottom == "moving" else u[:1]
        fu[:1] = -wall_shear_excess(us, y, ops.nu) * inv_h
    if ops.bc_top == "noslip":
        fu[-1:] = -wall_shear_excess(u[-1:], y, ops.nu) * inv_h
    if not ops.closed or (ops.bc_left != "noslip" and ops.bc_right != "noslip"):
        return fu, None
    fv = torch.zeros_like(v)
    if ops.bc_left == "noslip":
        fv[:, :1] = -wall_shear_excess(v[:, :1], y, ops.nu) * inv_h
    if ops.bc_right == "noslip":
        fv[:, -1:] = -wall_shear_excess(v[:, -1:], y, ops.nu) * inv_h
    return fu, fv


def momentum_rhs(u, v, geom, ops):
    """Tendencies with boundary rows set, max nu_t, nu_t field (core.py:571-614)."""
    U, V = pad_u(u, ops), pad_v(v, ops)
    au, av = advection(U, V, ops)
    nu_h2 = ops.nu * (1.0 / ops.h**2)
    lu = add_diffusion(au, U, nu_h2)
    lv = add_diffusion(av, V, nu_h2)
    nut_max = nut = None
    if ops.turbulence != "none":
        fu, fv, nut_max, nut = sgs_stress_divergence(U, V, geom, ops)
        lu = lu + fu
        lv = lv + fv
    if ops.wall_model:
        wu, wv = wall_model_force(u, v, ops)
        lu = lu + wu
        if wv is not None:
            lv = lv + wv
    if ops.periodic:
        return lu, lv, nut_max, nut
    if ops.closed:
        out = torch.zeros_like(u[:, :1])
    else:
        out = (u[:, -2:-1] - u[:, -1:]) * (1.0 / ops.h)
    lu = torch.cat([torch.zeros_like(out), lu[:, 1:-1], out], dim=1)
    zero = torch.zeros_like(lv[:1])
    lv = torch.cat([zero, lv[1:-1], zero], dim=0)
    return lu, lv, nut_max, nut


# ---------------------------------------------------------------------------
# projection, penalisation, dt (core.py:621-695)
# ---------------------------------------------------------------------------

def divergence(u, v, h):
    return (u[:, 1:] - u[:, :-1] + v[1:] - v[:-1]) * (1.0 / h)


def correct(u, v, psi, ops):
    """u - grad psi, boundary-normal faces untouched (core.py:645-660)."""
    inv_h = 1.0 / ops.h
    if ops.periodic:
        gx = (psi - torch.roll(psi, 1, dims=1)) * inv_h
        gy = (psi - torch.roll(psi, 1, dims=0)) * inv_h
        return u - torch.cat([gx, gx[:, :1]], dim=1), v - torch.cat([gy, gy[:1]], dim=0)
    gx = (psi[:, 1:] - psi[:, :-1]) * inv_h
    gy = (psi[1:] - psi[:-1]) * inv_h
    zc = torch.zeros_like(psi[:, :1])
    zr = torch.zeros_like(psi[:1])
    return u - torch.cat([zc, gx, zc], dim=1), v - torch.cat([zr, gy, zr], dim=0)


def project(u, v, ops):
    """(u, v, psi, rhs) with div = 0 to round-off (core.py:626-642)."""
    rhs = divergence(u, v, ops.h)
    psi = poisson(ops).solve(rhs)
    un, vn = correct(u, v, psi, ops)
    return un, vn, psi, rhs


def penalise(u, v, geom, ops, bdt):
    """Implicit Brinkman on the window; returns increments (core.py:663-682)."""
    j0, j1, i0, i1 = geom.window
    k = bdt / ops.eta
    u = u.clone()
    v = v.clone()
    uw = u[j0:j1, i0:i1 + 1]
    vw = v[j0:j1 + 1, i0:i1]
    un = (uw + k * geom.a_u) / (1.0 + k * geom.chi_u)
    vn = (vw + k * geom.a_v) / (1.0 + k * geom.chi_v)
    du, dv = un - uw, vn - vw
    u[j0:j1, i0:i1 + 1] = un
    v[j0:j1 + 1, i0:i1] = vn
    return u, v, du, dv


def cfl_dt(u, v, ops, nut_max=None):
    """min(cfl h / (max|u| + max|v|), 0.25 h^2 / (nu + max nu_t)) (core.py:685-695)."""
    umax = (u.abs().amax() + v.abs().amax()).clamp_min(1e-6)
    dt = ops.cfl * ops.h / umax
    nu = ops.nu if nut_max is None else ops.nu + nut_max
    return torch.minimum(dt, 0.25 * ops.h * ops.h / nu)


def physics_dt_limit(dt, u, v, theta, nut_max, ph, ops):
    """Thermal positivity and free-fall limits, heater limits (physics.py:851-888)."""
    if ph.thermal and theta is not None:
        kap = ph.kappa if nut_max is None else ph.kappa + nut_max * (1.0 / ph.prt)
        speed = torch.maximum(u.abs().amax() + v.abs().amax(), ph.u_floor)
        rate = speed * (2.0 / ops.h) + kap * (ph.diff_coef / (ops.h * ops.h))
        dt = torch.minimum(dt, 1.0 / rate)
        acc = ph.ri.abs() * theta.abs().amax()
        dt = torch.minimum(dt, ops.cfl * torch.sqrt(ops.h / acc.clamp_min(1e-12)))
        if ph.heat_max is not None:
            s_max = ph.heat_max.clamp_min(1e-12)
            dt = torch.minimum(dt, 1.0 / s_max)
            dt = torch.minimum(dt, (6.0 * ops.cfl * ops.h / (ph.ri.abs() * s_max).clamp_min(1e-12)) ** (1.0 / 3.0))
    return dt


PID_TARGET = 0.9
PID_GAINS = (0.10, 0.25, 0.02)    # (kP, kI, kD)
PID_MAX_GROWTH = 1.03


def pid_dt(dt_limit, dt_prev, ctrl):
    """PID-controlled step in log space (diagnostics.py:104-123): returns (dt, new ctrl)."""
    kp, ki, kd = PID_GAINS
    target = PID_TARGET * dt_limit
    first = dt_prev <= 0
    safe_prev = torch.where(first, target, dt_prev)
    e = torch.log(target / safe_prev)
    e1, e2 = ctrl[0], ctrl[1]
    step = ki * e + kp * (e - e1) + kd * (e - 2.0 * e1 + e2)
    grow = torch.exp(step.clamp(max=math.log(PID_MAX_GROWTH)))
    dt = torch.where(first, target, torch.minimum(safe_prev * grow, dt_limit))
    e = torch.where(first, torch.zeros_like(e), e)
    return dt, torch.cat([torch.stack([e, e1]), ctrl[2:]])


# ---------------------------------------------------------------------------
# scalars (scalar.py:35-141, core.py:702-771, physics.py:623-830)
# ---------------------------------------------------------------------------

def van_leer(a, b):
    ab = a * b
    s = a + b
    return torch.where(ab > 0, 2.0 * ab / torch.where(ab > 0, s, torch.ones_like(s)), torch.zeros_like(ab))


def _ghost_x(c, ops, inflow):
    if ops.periodic:
        return torch.cat([c[..., -2:], c, c[..., :2]], dim=-1)
    right = c[..., -1:].expand(*c.shape[:-1], 2)
    if ops.closed or inflow is None:
        left = c[..., :1].expand(*c.shape[:-1], 2)
    else:
        left = inflow.expand(*c.shape[:-1], 2) if torch.is_tensor(inflow) else torch.full_like(right, float(inflow))
    return torch.cat([left, c, right], dim=-1)


def _ghost_y(c, ops):
    if ops.periodic:
        return torch.cat([c[..., -2:, :], c, c[..., :2, :]], dim=-2)
    shape = list(c.shape)
    shape[-2] = 2
    return torch.cat([c[..., :1, :].expand(shape), c, c[..., -1:, :].expand(shape)], dim=-2)


def tvd_fluxes(c, u, v, ops, inflow=None):
    """MUSCL / van Leer face fluxes (scalar.py:63-83)."""
    cx = _ghost_x(c, ops, inflow)
    s = van_leer(cx[..., 1:-1] - cx[..., :-2], cx[..., 2:] - cx[..., 1:-1])
    left = cx[..., 1:-2] + 0.5 * s[..., :-1]
    right = cx[..., 2:-1] - 0.5 * s[..., 1:]
    fx = u * torch.where(u > 0, left, right)
    cy = _ghost_y(c, ops)
    s = van_leer(cy[..., 1:-1, :] - cy[..., :-2, :], cy[..., 2:, :] - cy[..., 1:-1, :])
    low = cy[..., 1:-2, :] + 0.5 * s[..., :-1, :]
    high = cy[..., 2:-1, :] - 0.5 * s[..., 1:, :]
    fy = v * torch.where(v > 0, low, high)
    return fx, fy


def flux_divergence(fx, fy, h):
    return (fx[..., :-1] - fx[..., 1:] + fy[..., :-1, :] - fy[..., 1:, :]) * (1.0 / h)


def diffusive_fluxes(c, kx, ky, ops, modes, values):
    """-kappa dc/dn at every face with wall conditions (scalar.py:99-141)."""
    h = ops.h
    inv_h, inv_half = 1.0 / h, 2.0 / h
    kx_t, ky_t = torch.is_tensor(kx) and kx.dim() > 0, torch.is_tensor(ky) and ky.dim() > 0
    kxi = kx[..., 1:-1] if kx_t else kx
    kyi = ky[..., 1:-1, :] if ky_t else ky
    fxi = -(kxi * (c[..., 1:] - c[..., :-1])) * inv_h
    fyi = -(kyi * (c[..., 1:, :] - c[..., :-1, :])) * inv_h
    kl = kx[..., :1] if kx_t else kx
    kr = kx[..., -1:] if kx_t else kx
    kb = ky[..., :1, :] if ky_t else ky
    kt = ky[..., -1:, :] if ky_t else ky
    first_x, last_x = c[..., :1], c[..., -1:]
    first_y, last_y = c[..., :1, :], c[..., -1:, :]

    def side(mode, k, inner, value, sign, wrap):
        if mode in ("temperature", "inlet"):
            return (-sign * inv_half) * k * (inner - value)
        if mode == "heat_flux":
            return (sign * k) * value * torch.ones_like(inner)
        if mode == "periodic":
            return wrap
        return torch.zeros_like(inner)

    wrap_x = -(kl * (first_x - last_x)) * inv_h if modes[0] == "periodic" else None
    wrap_y = -(kb * (first_y - last_y)) * inv_h if modes[2] == "periodic" else None
    fl = side(modes[0], kl, first_x, values[0], +1.0, wrap_x)
    fr = side(modes[1], kr, last_x, values[1], -1.0, wrap_x)
    fb = side(modes[2], kb, first_y, values[2], +1.0, wrap_y)
    ft = side(modes[3], kt, last_y, values[3], -1.0, wrap_y)
    return torch.cat([fl, fxi, fr], dim=-1), torch.cat([fb, fyi, ft], dim=-2)


def face_average_x(q, ops):
    mid = 0.5 * (q[..., 1:] + q[..., :-1])
    if ops.periodic:
        w = 0.5 * (q[..., :1] + q[..., -1:])
        return torch.cat([w, mid, w], dim=-1)
    return torch.cat([q[..., :1], mid, q[..., -1:]], dim=-1)


def face_average_y(q, ops):
    mid = 0.5 * (q[..., 1:, :] + q[..., :-1, :])
    if ops.periodic:
        w = 0.5 * (q[..., :1, :] + q[..., -1:, :])
        return torch.cat([w, mid, w], dim=-2)
    return torch.cat([q[..., :1, :], mid, q[..., -1:, :]], dim=-2)


def _window_apply(fld, win, fn, extra_cols=0, extra_rows=0):
    j0, j1, i0, i1 = win
    part = fn(fld[..., j0:j1 + extra_rows, i0:i1 + extra_cols])
    out = fld.clone()
    out[..., j0:j1 + extra_rows, i0:i1 + extra_cols] = part
    return out


def theta_rhs(th, u, v, nut, geom, ph, ops):
    """Temperature tendency: TVD advection, diffusion (1/Pe + nu_t/Pr_t, insulated solids,
    wall conditions), heaters (physics.py:696-730)."""
    fx, fy = tvd_fluxes(th, u, v, ops, inflow=ph.theta_in)
    lt = flux_divergence(fx, fy, ops.h)
    tg = geom.thermal if geom is not None else None
    insulated = tg is not None and not geom.empty
    if nut is not None:
        nc = nut[1:-1, 1:-1] * (1.0 / ph.prt)
        kx = ph.kappa + face_average_x(nc, ops)
        ky = ph.kappa + face_average_y(nc, ops)
    elif insulated:
        kx = ph.kappa.expand(ops.ny, ops.nx + 1)
        ky = ph.kappa.expand(ops.ny + 1, ops.nx)
    else:
        kx = ky = ph.kappa
    if insulated:
        kx = _window_apply(kx.contiguous(), geom.window, lambda b: b * tg.ins_u, extra_cols=1)
        ky = _window_apply(ky.contiguous(), geom.window, lambda b: b * tg.ins_v, extra_rows=1)
    Fx, Fy = diffusive_fluxes(th, kx, ky, ops, ph.wall_modes, ph.wall_values)
    lt = lt + flux_divergence(Fx, Fy, ops.h)
    if ph.heat is not None:
        lt = lt + ph.heat
    if tg is not None and not geom.empty and tg.flux_ids:
        lt = _window_apply(lt, geom.window, lambda b: b + ph.kappa * tg.q_skin)
    return lt


def penalise_theta(th, geom, ph, bdt):
    """Implicit penalisation toward the temperature bodies; per-body heat (physics.py:733-757)."""
    k = bdt / ph.eta_t
    heat = None
    tg = geom.thermal if geom is not None else None
    if tg is not None and not geom.empty and tg.temp_ids:
        j0, j1, i0, i1 = geom.window
        tw = th[j0:j1, i0:i1]
        tn = (tw + k * tg.a_t) / (1.0 + k * tg.chi_t)
        heat = torch.einsum("bij,ij->b", tg.w_t, tn - tw)
        th = th.clone()
        th[j0:j1, i0:i1] = tn
    return th, heat


def buoyancy(theta, ph, ops):
    """-Ri theta g on the faces, zero on boundary faces (physics.py:623-637)."""
    gx, gy = ph.gravity
    fu = fv = None
    if gy != 0.0:
        tv = face_average_y(theta, ops)
        if not ops.periodic:
            tv = torch.cat([torch.zeros_like(tv[:1]), tv[1:-1], torch.zeros_like(tv[:1])], dim=0)
        fv = tv * (ph.ri * (-gy))
    if gx != 0.0:
        tu = face_average_x(theta, ops)
        if not ops.periodic:
            tu = torch.cat([torch.zeros_like(tu[:, :1]), tu[:, 1:-1], torch.zeros_like(tu[:, :1])], dim=1)
        fu = tu * (ph.ri * (-gx))
    return fu, fv


def add_buoyancy(uh, vh, theta, geom, ops, ph, bdt, w, acc_u, acc_v):
    """Boussinesq impulse after penalisation, solid share booked to the force (core.py:774-812)."""
    fu, fv = buoyancy(theta, ph, ops)
    if fu is not None:
        uh = uh + bdt * fu
    if fv is not None:
        vh = vh + bdt * fv
    if not geom.empty:
        j0, j1, i0, i1 = geom.window
        tg = geom.thermal
        vu = tg.vol_u if tg is not None and tg.vol_u is not None else geom.chi_u
        vv = tg.vol_v if tg is not None and tg.vol_v is not None else geom.chi_v
        if fu is not None:
            share = fu[j0:j1, i0:i1 + 1] * vu * (bdt * w)
            acc_u = share if acc_u is None else acc_u + share
        if fv is not None:
            share = fv[j0:j1 + 1, i0:i1] * vv * (bdt * w)
            acc_v = share if acc_v is None else acc_v + share
    return uh, vh, acc_u, acc_v


def emitter_on(t, timing):
    start, stop, period, duty = timing[0], timing[1], timing[2], timing[3]
    active = (t >= start) & (t < stop)
    safe = torch.where(period > 0, period, torch.ones_like(period))
    phase = torch.remainder(t - start, safe) / safe
    on = active & ((period <= 0) | (phase < duty))
    return on.to(t.dtype)


def update_dye(c, u, v, dt, t, geom, ph, ops):
    """Forward-Euler TVD smoke step, emitters, solids, floor and clip (physics.py:805-830)."""
    fx, fy = tvd_fluxes(c, u, v, ops, inflow=ph.dye_inflow)
    c = c + dt * flux_divergence(fx, fy, ops.h)
    if ph.groups:
        chans = list(c.unbind(0))
        for g in ph.groups:
            wgt = g.weight * emitter_on(t, g.timing)
            chans[g.channel] = torch.addcmul(chans[g.channel], wgt, g.target - chans[g.channel])
        c = torch.stack(chans)
    if geom is not None and not geom.empty:
        keep = (geom.chi_c <= 0.5).to(c.dtype)
        c = _window_apply(c, geom.window, lambda b: b * keep)
    return torch.nn.functional.threshold(c, DYE_FLOOR, 0.0).clamp(max=1.0)


def advect_dye(c, u, v, dt, ops):
    """Inlet-comb smoke, forward Euler MUSCL (core.py:709-733)."""
    if ops.periodic:
        cx = torch.cat([c[:, -2:], c, c[:, :2]], dim=1)
        cy = torch.cat([c[-2:], c, c[:2]], dim=0)
    else:
        cx = torch.cat([ops.dye_inlet.expand(-1, 2), c, c[:, -1:].expand(-1, 2)], dim=1)
        cy = torch.cat([c[:1].expand(2, -1), c, c[-1:].expand(2, -1)], dim=0)
    s = van_leer(cx[:, 1:-1] - cx[:, :-2], cx[:, 2:] - cx[:, 1:-1])
    left = cx[:, 1:-2] + 0.5 * s[:, :-1]
    right = cx[:, 2:-1] - 0.5 * s[:, 1:]
    fx = u * torch.where(u > 0, left, right)
    s = van_leer(cy[1:-1] - cy[:-2], cy[2:] - cy[1:-1])
    low = cy[1:-2] + 0.5 * s[:-1]
    high = cy[2:-1] - 0.5 * s[1:]
    fy = v * torch.where(v > 0, low, high)
    c = c - (dt * (1.0 / ops.h)) * (fx[:, 1:] - fx[:, :-1] + fy[1:] - fy[:-1])
    return torch.nn.functional.threshold(c, 1e-6, 0.0).clamp(max=1.0)


def zero_dye_in_solids(dye, geom):
    if geom.empty:
        return dye
    j0, j1, i0, i1 = geom.window
    dye = dye.clone()
    dye[..., j0:j1, i0:i1].mul_((geom.chi_c <= 0.5).to(dye.dtype))
    return dye


def wall_heat_rates(th, ph, ops):
    """(left, right, bottom, top) conductive heat into the fluid (physics.py:760-781)."""
    k, h = ph.kappa, ops.h
    first = (th[:, 0], th[:, -1], th[0, :], th[-1, :])
    rates = []
    for i, mode in enumerate(ph.wall_modes):
        if mode in ("temperature", "inlet"):
            rates.append((2.0 * k) * (ph.wall_values[i] - first[i]).sum())
        elif mode == "heat_flux":
            rates.append(k * ph.wall_values[i] * (first[i].numel() * h))
        else:
            rates.append(torch.zeros((), dtype=th.dtype))
    return torch.stack(rates)


def flow_heat_rates(th, u, ph, ops):
    """Advective heat in at the inlet and out at the outlet (physics.py:784-795)."""
    if ops.periodic or ops.closed:
        return torch.zeros(2, dtype=th.dtype)
    h = ops.h
    return torch.stack([(u[:, 0] * ph.theta_in).sum() * h, (u[:, -1] * th[:, -1]).sum() * h])


# ---------------------------------------------------------------------------
# the step (core.py:815-969)
# ---------------------------------------------------------------------------

def change_norms(new, old, h):
    """[L2, RMS, max, relative L2] of the change of one field over a step, in float64
    (diagnostics.py:55-68)."""
    d = (new - old).to(torch.float64)
    nrm = math.sqrt(float((d * d).sum()))
    l2 = nrm * h
    ref = math.sqrt(float((new.to(torch.float64) ** 2).sum())) * h
    return [l2, nrm / math.sqrt(d.numel()), float(d.abs().max()), l2 / max(ref, 1e-300)]


def diagnostics_row(old, new, psi, rhs, ops):
    """One diagnostics row without "step" (22 float64 values) from the states before and
    after a step and the last stage's potential and Poisson right-hand side
    (diagnostics.py:70-101): t, dt, CFL = dt (max|u| + max|v|) / h, RMS and max of the new
    state's divergence, ||L psi - f|| / ||f||, then the change norms of u, v, p, theta (NaN
    without temperature)."""
    h = ops.h
    div = divergence(new.u, new.v, h).to(torch.float64)
    dt = float(new.dt)
    umax = float(new.u.abs().max()) + float(new.v.abs().max())
    r = (poisson(ops).laplacian(psi) - rhs).to(torch.float64)
    f2 = float((rhs.to(torch.float64) ** 2).sum())
    row = [float(new.t), dt, dt * umax / h, math.sqrt(float((div * div).sum())) / math.sqrt(div.numel()),
           float(div.abs().max()), math.sqrt(float((r * r).sum())) / max(math.sqrt(f2), 1e-30)]
    row += change_norms(new.u, old.u, h) + change_norms(new.v, old.v, h) + change_norms(new.p, old.p, h)
    if new.theta is not None and old.theta is not None:
        row += change_norms(new.theta, old.theta, h)
    else:
        row += [math.nan] * 4
    return torch.tensor(row, dtype=torch.float64)


def step_full(state: State, geom, ops, dt=None, *, out: dict | None = None, trace: dict | None = None,
              diag: bool = False):
    """One SSP-RK3 step; returns (new State, coeffs (nbody, 3)) (core.py:815-969). With
    ``diag`` and ``out``, ``out["diag"]`` is the step's diagnostics row (core.py:923-925, 1003-1004)."""
    ph = ops.phys
    u0, v0 = state.u, state.v
    up, vp = u0, v0
    thermal = ph is not None and ph.thermal and state.theta is not None
    th0 = state.theta if thermal else None
    thp = th0
    acc_u = acc_v = None
    psi_w = None
    heat_acc = None
    want_nut = thermal and ops.turbulence != "none"
    ctrl = state.ctrl
    for s, ((a, b), w) in enumerate(zip(SSP_RK3, FORCE_WEIGHTS)):
        lu, lv, nut_max, nut = momentum_rhs(up, vp, geom, ops)
        if s == 0 and dt is None:
            if ops.dt_fixed is not None:
                dt = ops.extras["dt_fixed"]
            else:
                dt = cfl_dt(u0, v0, ops, nut_max)
                if ph is not None:
                    dt = physics_dt_limit(dt, u0, v0, th0, nut_max, ph, ops)
                    if ph.pid:
                        dt, ctrl = pid_dt(dt, state.dt, ctrl)
        if not torch.is_tensor(dt):
            dt = torch.tensor(float(dt), dtype=ops.dtype)
        lt = theta_rhs(thp, up, vp, nut if want_nut else None, geom, ph, ops) if thermal else None
        bdt = dt * b
        if a == 0.0:
            uh = torch.addcmul(up, lu, bdt)
            vh = torch.addcmul(vp, lv, bdt)
        else:
            uh = torch.addcmul(torch.lerp(up, u0, a), lu, bdt)
            vh = torch.addcmul(torch.lerp(vp, v0, a), lv, bdt)
        if lt is not None:
            thh = torch.addcmul(thp if a == 0.0 else torch.lerp(thp, th0, a), lt, bdt)
        if not ops.periodic:
            uh = inlet_outlet(uh, ops)
        if trace is not None:
            trace[f"stage{s + 1}/explicit_u"], trace[f"stage{s + 1}/explicit_v"] = uh.clone(), vh.clone()
        if not geom.empty:
            uh, vh, du, dv = penalise(uh, vh, geom, ops, bdt)
            acc_u = du * w if acc_u is None else acc_u + du * w
            acc_v = dv * w if acc_v is None else acc_v + dv * w
        if thermal and ph.buoyant:
            uh, vh, acc_u, acc_v = add_buoyancy(uh, vh, thp, geom, ops, ph, bdt, w, acc_u, acc_v)
        if ops.periodic and not geom.empty:
            uh, vh = apply_boundary_values(uh, vh, ops)
        if lt is not None:
            thh, heat = penalise_theta(thh, geom, ph, bdt)
            if heat is not None:
                heat_acc = heat * w if heat_acc is None else heat_acc + heat * w
            thp = thh
        if trace is not None:
            trace[f"stage{s + 1}/penalised_u"], trace[f"stage{s + 1}/penalised_v"] = uh.clone(), vh.clone()
        up, vp, psi, rhs = project(uh, vh, ops)
        if trace is not None:
            trace[f"stage{s + 1}/divergence"] = rhs
            trace[f"stage{s + 1}/pressure_potential"] = psi
            trace[f"stage{s + 1}/projected_u"], trace[f"stage{s + 1}/projected_v"] = up, vp
        psi_w = psi * w if psi_w is None else torch.add(psi_w, psi, alpha=w)
    p = psi_w * (1.0 / dt)
    if geom.empty:
        coeffs = torch.zeros(geom.nbody, 3, dtype=ops.dtype)
    else:
        scale = -(ops.h * ops.h) / (dt * 0.5 * ops.force_ref)
        ru = torch.einsum("bkij,ij->bk", geom.k_u, acc_u)
        rv = torch.einsum("bkij,ij->bk", geom.k_v, acc_v)
        coeffs = torch.stack([ru[:, 0], rv[:, 0], ru[:, 1] + rv[:, 1]], dim=1) * scale
    dye, dye_extra = state.dye, state.dye_extra
    t_new = state.t + dt
    if ph is not None and ph.n_dye:
        c = dye[None] if dye_extra is None else torch.cat([dye[None], dye_extra])
        c = update_dye(c, up, vp, dt, t_new, geom, ph, ops)
        dye, dye_extra = c[0], (c[1:] if c.shape[0] > 1 else None)
    elif dye is not None:
        dye = zero_dye_in_solids(advect_dye(dye, up, vp, dt, ops), geom)
    new = State(up, vp, p, dye, t_new, dt * torch.ones_like(state.t), thp, dye_extra, ctrl)
    if out is not None and diag:
        out["diag"] = diagnostics_row(state, new, psi, rhs, ops)
    if out is not None and thermal:
        h2 = ops.h * ops.h
        tg = geom.thermal
        if tg is not None and tg.temp_ids:
            out["heat"] = heat_acc * (h2 / dt)
        if tg is not None and tg.flux_ids and not geom.empty:
            # mean skin temperature of each heat-flux body (core.py:988-990)
            j0, j1, i0, i1 = geom.window
            out["theta_wall"] = torch.einsum("bij,ij->b", tg.w_f, new.theta[j0:j1, i0:i1])
        out["wall_heat"] = wall_heat_rates(new.theta, ph, ops)
        out["heat_storage"] = ((new.theta.sum() - state.theta.sum()) * (h2 / dt)).reshape(1)
        out["flow_heat"] = flow_heat_rates(new.theta, new.u, ph, ops)
    return new, coeffs


def init_state(cfg, ops, perturb=None) -> State:
    """Impulsive start (core.py:216-271): the inlet profile everywhere with v = 0 (zero flow
    in a periodic or closed box); with ``perturb`` and bodies in a tunnel, a Lamb-Oseen
    vortex of peak swirl 0.05 centred on the bodies' bounding box (core radius 0.75 of its
    size), boundary values and projection. Smoke and temperature start as the reference's."""
    import math

    import numpy as np
    perturb = cfg.perturb if perturb is None else perturb
    dt_ = ops.dtype
    tunnel = not (ops.periodic or ops.closed)
    if tunnel:
        u = ops.u_inlet.expand(ops.ny, ops.nx + 1).clone()
    else:
        u = torch.zeros(ops.ny, ops.nx + 1, dtype=dt_)
    v = torch.zeros(ops.ny + 1, ops.nx, dtype=dt_)
    if perturb and cfg.bodies and tunnel:
        pts = np.vstack([b.points for b in cfg.bodies])
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        cx, cy = 0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1])
        size = max(hi[1] - lo[1], 0.25 * (hi[0] - lo[0]), 4 * ops.h)
        rc = 0.75 * size
        gam = 0.05 * 2.0 * math.pi * rc / 0.638
        xc = ops.x0 + (torch.arange(ops.nx, dtype=dt_) + 0.5) * ops.h
        yc = ops.y0 + (torch.arange(ops.ny, dtype=dt_) + 0.5) * ops.h
        xu = ops.x0 + torch.arange(ops.nx + 1, dtype=dt_) * ops.h
        yv = ops.y0 + torch.arange(ops.ny + 1, dtype=dt_) * ops.h

        def swirl(x, y):
            dx, dy = x - cx, y - cy
            r2 = (dx * dx + dy * dy).clamp_min(1e-12)
            f = gam / (2 * math.pi * r2) * (1 - torch.exp(-r2 / rc**2))
            return -f * dy, f * dx
        yy, xx = torch.meshgrid(yc, xu, indexing="ij")
        u = u + swirl(xx, yy)[0]
        yy, xx = torch.meshgrid(yv, xc, indexing="ij")
        v = v + swirl(xx, yy)[1]
        u, v = apply_boundary_values(u, v, ops)
        u, v, _, _ = project(u, v, ops)
    zero = torch.zeros((), dtype=dt_)
    ph = ops.phys
    n_dye = cfg.dye_channels
    dye = torch.zeros(ops.ny, ops.nx, dtype=dt_) if n_dye else None
    theta = dye_extra = ctrl = None
    if ph is not None:
        if ph.thermal:
            th_init = ph.theta_init if not tunnel else float(cfg.thermal.theta_inlet)
            theta = torch.full((ops.ny, ops.nx), th_init, dtype=dt_)
        if n_dye > 1:
            dye_extra = torch.zeros(n_dye - 1, ops.ny, ops.nx, dtype=dt_)
        if getattr(ph, "ctrl_size", 0):
            ctrl = torch.zeros(ph.ctrl_size, dtype=dt_)
    return State(u, v, torch.zeros(ops.ny, ops.nx, dtype=dt_), dye, zero.clone(), zero.clone(), theta, dye_extra, ctrl)


def noisy_start(ops, seed: int, amplitude: float = 0.01):
    """Inlet flow + seeded Gaussian noise on every face, boundary values, projection
    (the noisy initial state of the benchmark cases; core.py:328-334, 626-642)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    nu_ = rng.standard_normal((ops.ny, ops.nx + 1))
    nv_ = rng.standard_normal((ops.ny + 1, ops.nx))
    base = 0.0 if (ops.periodic or ops.closed) else 1.0
    u = torch.as_tensor(base + amplitude * nu_, dtype=ops.dtype)
    v = torch.as_tensor(amplitude * nv_, dtype=ops.dtype)
    u, v = apply_boundary_values(u, v, ops)
    u, v, _, _ = project(u, v, ops)
    return u, v
