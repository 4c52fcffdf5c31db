"""CPU oracle of the 3D Navier-Stokes step (TEST INFRASTRUCTURE, never the product).

A restatement, in plain torch on the CPU (float32 / float64), of the reference's 3D step
``step3d_full`` (``solvers/ns3d/core.py:753-883``) and its helpers, without placed physics:
ghost rules per component and axis, upwind3 / central advection, 7-point diffusion,
Smagorinsky / WALE / Vreman eddy viscosity on the traceless gradient, window penalisation,
the exact Poisson solve with orthonormal DCT-II / Hartley bases (``solvers/ns3d/poisson.py``),
the step-averaged pressure and the six force / moment coefficients.

Only ``tests/`` and ``bench.py``'s CPU-baseline leg use it, as the checker. It takes plain
data objects (``ops`` with the fields of the reference's ``Ops3D``, ``geom`` with those of
``Geometry3D``).

Pinning: the reference's own trajectories (``tests/golden/make_ns3d_fixtures.py``), the
Poisson residual and round trip (closed form), divergence of projected noise, Taylor-Green
decay, zero eddy viscosity in pure shear for WALE / Vreman.
"""
from __future__ import annotations

import math
from typing import NamedTuple

import torch
import torch.nn.functional as F

SSP_RK3 = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))
FORCE_WEIGHTS = (1.0 / 6.0, 2.0 / 3.0, 1.0)
VISCOUS_DT = 1.0 / 6.0                                   # ns3d/core.py:51
SQRT_EPS = 1e-30


class State3D(NamedTuple):
    u: torch.Tensor
    v: torch.Tensor
    w: torch.Tensor
    p: torch.Tensor
    t: torch.Tensor
    dt: torch.Tensor


# ---------------------------------------------------------------------------
# Poisson (ns3d/poisson.py:74-157)
# ---------------------------------------------------------------------------

def neumann_basis(n):
    k = torch.arange(n, dtype=torch.float64)[:, None]
    j = torch.arange(n, dtype=torch.float64)[None, :]
    q = torch.cos(math.pi * k * (j + 0.5) / n) * math.sqrt(2.0 / n)
    q[0] *= math.sqrt(0.5)
    lam = 2.0 * torch.cos(math.pi * torch.arange(n, dtype=torch.float64) / n) - 2.0
    return q, lam


def periodic_basis(n):
    k = torch.arange(n, dtype=torch.float64)
    ang = 2.0 * math.pi * k[:, None] * k[None, :] / n
    q = (torch.cos(ang) + torch.sin(ang)) / math.sqrt(n)
    lam = 2.0 * torch.cos(2.0 * math.pi * k / n) - 2.0
    return q, lam


class PoissonSolver3D:
    """Exact solve with orthonormal per-axis bases applied as matrix products."""

    def __init__(self, nx, ny, nz, h, *, periodic, dtype):
        self.nx, self.ny, self.nz, self.h = nx, ny, nz, h
        self.periodic = tuple(bool(p) for p in periodic)
        mats, lams = [], []
        for n, per in zip((nx, ny, nz), self.periodic):
            q, lam = periodic_basis(n) if per else neumann_basis(n)
            mats.append(q)
            lams.append(lam / (h * h))
        self.qx, self.qy, self.qz = (m.to(dtype) for m in mats)
        lx, ly, lz = lams
        lam = lz[:, None, None] + ly[None, :, None] + lx[None, None, :]
        inv = torch.where(lam != 0, 1.0 / torch.where(lam != 0, lam, torch.ones_like(lam)), torch.zeros_like(lam))
        inv[0, 0, 0] = 0.0
        self.inv_lam = inv.to(dtype)

    def solve(self, f):
        nz, ny, nx = f.shape
        c = (f.reshape(-1, nx) @ self.qx.t()).reshape(nz, ny, nx)
        c = torch.matmul(self.qy, c)
        c = (self.qz @ c.reshape(nz, -1)).reshape(nz, ny, nx)
        c = c * self.inv_lam
        c = (self.qz.t() @ c.reshape(nz, -1)).reshape(nz, ny, nx)
        c = torch.matmul(self.qy.t(), c)
        return (c.reshape(-1, nx) @ self.qx).reshape(nz, ny, nx)

    def laplacian(self, psi):
        pe = psi
        for dim, per in zip((2, 1, 0), self.periodic):
            n = pe.shape[dim]
            if per:
                lo, hi = pe.narrow(dim, n - 1, 1), pe.narrow(dim, 0, 1)
            else:
                lo, hi = pe.narrow(dim, 0, 1), pe.narrow(dim, n - 1, 1)
            pe = torch.cat([lo, pe, hi], dim=dim)
        c = pe[1:-1, 1:-1, 1:-1]
        ring = (pe[1:-1, 1:-1, 2:] + pe[1:-1, 1:-1, :-2] + pe[1:-1, 2:, 1:-1] + pe[1:-1, :-2, 1:-1]
                + pe[2:, 1:-1, 1:-1] + pe[:-2, 1:-1, 1:-1])
        return (ring - 6.0 * c) / (self.h * self.h)


_POISSON: dict = {}


def poisson(ops):
    key = (ops.nx, ops.ny, ops.nz, ops.h, tuple(ops.periodic_axes), ops.dtype)
    if key not in _POISSON:
        _POISSON[key] = PoissonSolver3D(ops.nx, ops.ny, ops.nz, ops.h, periodic=ops.periodic_axes, dtype=ops.dtype)
    return _POISSON[key]


# ---------------------------------------------------------------------------
# ghosts (ns3d/core.py:147-278)
# ---------------------------------------------------------------------------

def ghost_rules(ops):
    """{comp: (x, y, z)} rules, each a (lo, hi) pair of (kind, value) (core.py:147-181)."""
    px, py, pz = ops.periodic_axes
    pc, pf = ("periodic_cell", 0.0), ("periodic_face", 0.0)
    even, odd, edge, fodd = ("even", 0.0), ("odd", 0.0), ("edge", 0.0), ("face_odd", 0.0)
    zlo_tan = {"slip": even, "noslip": odd, "moving": ("affine", float(ops.ground_speed))}[ops.bc_zmin]
    zlo_v = {"slip": even, "noslip": odd, "moving": odd}[ops.bc_zmin]
    zhi = {"slip": even, "noslip": odd}[ops.bc_zmax]
    y_tan = odd if ops.bc_y == "noslip" else even
    if ops.closed:
        x_tan = odd if ops.bc_x == "noslip" else even
        ux, vx = (fodd, fodd), (x_tan, x_tan)
    else:
        ux, vx = (edge, edge), (odd, edge)
    u = ((pf, pf) if px else ux, (pc, pc) if py else (y_tan, y_tan), (pc, pc) if pz else (zlo_tan, zhi))
    v = ((pc, pc) if px else vx, (pf, pf) if py else (fodd, fodd), (pc, pc) if pz else (zlo_v, zhi))
    w = ((pc, pc) if px else vx, (pc, pc) if py else (y_tan, y_tan), (pf, pf) if pz else (fodd, fodd))
    return {"u": u, "v": v, "w": w}


def _ghost_slabs(x, dim, lo, hi):
    n = x.shape[dim]

    def make(rule, side):
        kind, val = rule
        if kind == "edge":
            return x.narrow(dim, 0 if side == "lo" else n - 1, 1)
        if kind in ("even", "odd", "affine"):
            m = x.narrow(dim, 0 if side == "lo" else n - 2, 2).flip(dim)
            if kind == "even":
                return m
            if kind == "odd":
                return -m
            return 2.0 * val - m
        if kind == "face_odd":
            return -x.narrow(dim, 1 if side == "lo" else n - 3, 2).flip(dim)
        if kind == "periodic_cell":
            return x.narrow(dim, n - 2 if side == "lo" else 0, 2)
        if kind == "periodic_face":
            return x.narrow(dim, n - 3 if side == "lo" else 1, 2)
        raise ValueError(kind)

    return make(lo, "lo"), make(hi, "hi")


def pad3(x, rules):
    """Two ghost layers on every side, axis by axis x, y, z (core.py:249-272)."""
    nz, ny, nx = x.shape
    out = x.new_empty(nz + 4, ny + 4, nx + 4)
    rx, ry, rz = rules
    lo, hi = _ghost_slabs(x, 2, *rx)
    out[2:-2, 2:-2, 2:-2] = x
    out[2:-2, 2:-2, 0:2] = lo
    out[2:-2, 2:-2, nx + 2:] = hi
    mid = out[2:-2]
    lo, hi = _ghost_slabs(mid[:, 2:-2], 1, *ry)
    out[2:-2, 0:2] = lo
    out[2:-2, ny + 2:] = hi
    lo, hi = _ghost_slabs(out[2:-2], 0, *rz)
    out[0:2] = lo
    out[nz + 2:] = hi
    return out


def pad_velocity(u, v, w, ops):
    g = ghost_rules(ops)
    return pad3(u, g["u"]), pad3(v, g["v"]), pad3(w, g["w"])


def inlet_outlet(u, ops):
    """Inlet plane from the profile, outlet shifted so outflow = inflow (core.py:281-297)."""
    if ops.closed:
        z = torch.zeros_like(u[..., :1])
        return torch.cat([z, u[..., 1:-1], z], dim=2)
    corr = (ops.u_inlet.sum() - u[..., -1].sum()) * (1.0 / (ops.nz * ops.ny))
    return torch.cat([ops.u_inlet, u[..., 1:-1], u[..., -1:] + corr], dim=2)


def sync_periodic(u, v, w, ops):
    px, py, pz = ops.periodic_axes
    if px:
        u = torch.cat([u[..., :-1], u[..., :1]], dim=2)
    if py:
        v = torch.cat([v[:, :-1], v[:, :1]], dim=1)
    if pz:
        w = torch.cat([w[:-1], w[:1]], dim=0)
    return u, v, w


def apply_boundary_values(u, v, w, ops):
    px, py, pz = ops.periodic_axes
    if not px:
        u = inlet_outlet(u, ops)
    if not py:
        z = torch.zeros_like(v[:, :1])
        v = torch.cat([z, v[:, 1:-1], z], dim=1)
    if not pz:
        z = torch.zeros_like(w[:1])
        w = torch.cat([z, w[1:-1], z], dim=0)
    return sync_periodic(u, v, w, ops)


# ---------------------------------------------------------------------------
# right-hand side (core.py:331-570)
# ---------------------------------------------------------------------------

def face_flux(vsum, a, b, c, d, upwind):
    s = b + c
    vsum = s if vsum is None else vsum
    if not upwind:
        return vsum * s
    cen = torch.add(s, a + d, alpha=-1.0 / 7.0)
    dis = torch.add(b - c, a - d, alpha=-1.0 / 3.0)
    return torch.addcmul(vsum * cen, vsum.abs(), dis, value=3.0 / 7.0)


def advection3(U, V, W, ops):
    """-div(u u) at u-, v-, w-faces (core.py:353-390)."""
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    up = ops.advection == "upwind3"
    k = (7.0 / 24.0 if up else 0.25) / ops.h
    m = slice(2, -2)
    fx = face_flux(None, U[m, m, 0:nx + 2], U[m, m, 1:nx + 3], U[m, m, 2:nx + 4], U[m, m, 3:nx + 5], up)
    au = fx[..., :-1] - fx[..., 1:]
    vs = V[m, 2:ny + 3, 1:nx + 2] + V[m, 2:ny + 3, 2:nx + 3]
    g = face_flux(vs, U[m, 0:ny + 1, m], U[m, 1:ny + 2, m], U[m, 2:ny + 3, m], U[m, 3:ny + 4, m], up)
    au = au + (g[:, :-1] - g[:, 1:])
    ws = W[2:nz + 3, m, 1:nx + 2] + W[2:nz + 3, m, 2:nx + 3]
    g = face_flux(ws, U[0:nz + 1, m, m], U[1:nz + 2, m, m], U[2:nz + 3, m, m], U[3:nz + 4, m, m], up)
    au = (au + (g[:-1] - g[1:])) * k
    us = U[m, 1:ny + 2, m] + U[m, 2:ny + 3, m]
    g = face_flux(us, V[m, m, 0:nx + 1], V[m, m, 1:nx + 2], V[m, m, 2:nx + 3], V[m, m, 3:nx + 4], up)
    av = g[..., :-1] - g[..., 1:]
    g = face_flux(None, V[m, 0:ny + 2, m], V[m, 1:ny + 3, m], V[m, 2:ny + 4, m], V[m, 3:ny + 5, m], up)
    av = av + (g[:, :-1] - g[:, 1:])
    ws = W[2:nz + 3, 1:ny + 2, m] + W[2:nz + 3, 2:ny + 3, m]
    g = face_flux(ws, V[0:nz + 1, m, m], V[1:nz + 2, m, m], V[2:nz + 3, m, m], V[3:nz + 4, m, m], up)
    av = (av + (g[:-1] - g[1:])) * k
    us = U[1:nz + 2, m, m] + U[2:nz + 3, m, m]
    g = face_flux(us, W[m, m, 0:nx + 1], W[m, m, 1:nx + 2], W[m, m, 2:nx + 3], W[m, m, 3:nx + 4], up)
    aw = g[..., :-1] - g[..., 1:]
    vs = V[1:nz + 2, m, m] + V[2:nz + 3, m, m]
    g = face_flux(vs, W[m, 0:ny + 1, m], W[m, 1:ny + 2, m], W[m, 2:ny + 3, m], W[m, 3:ny + 4, m], up)
    aw = aw + (g[:, :-1] - g[:, 1:])
    g = face_flux(None, W[0:nz + 2, m, m], W[1:nz + 3, m, m], W[2:nz + 4, m, m], W[3:nz + 5, m, m], up)
    aw = (aw + (g[:-1] - g[1:])) * k
    return au, av, aw


def laplacian_padded(P):
    m = slice(2, -2)
    ring = (P[m, m, 3:-1] + P[m, m, 1:-3]) + (P[m, 3:-1, m] + P[m, 1:-3, m]) + (P[3:-1, m, m] + P[1:-3, m, m])
    return ring - 6.0 * P[m, m, m]


def velocity_gradient(U, V, W, h):
    """Gradient at cells -1..n and edge shears (core.py:400-441)."""
    nz, ny, nx = U.shape[0] - 4, U.shape[1] - 4, U.shape[2] - 5
    inv = 1.0 / h
    g11 = (U[1:nz + 3, 1:ny + 3, 2:nx + 4] - U[1:nz + 3, 1:ny + 3, 1:nx + 3]) * inv
    g22 = (V[1:nz + 3, 2:ny + 4, 1:nx + 3] - V[1:nz + 3, 1:ny + 3, 1:nx + 3]) * inv
    g33 = (W[2:nz + 4, 1:ny + 3, 1:nx + 3] - W[1:nz + 3, 1:ny + 3, 1:nx + 3]) * inv
    dudy = (U[1:nz + 3, 1:ny + 4, 1:nx + 4] - U[1:nz + 3, 0:ny + 3, 1:nx + 4]) * inv
    dvdx = (V[1:nz + 3, 1:ny + 4, 1:nx + 4] - V[1:nz + 3, 1:ny + 4, 0:nx + 3]) * inv
    dudz = (U[1:nz + 4, 1:ny + 3, 1:nx + 4] - U[0:nz + 3, 1:ny + 3, 1:nx + 4]) * inv
    dwdx = (W[1:nz + 4, 1:ny + 3, 1:nx + 4] - W[1:nz + 4, 1:ny + 3, 0:nx + 3]) * inv
    dvdz = (V[1:nz + 4, 1:ny + 4, 1:nx + 3] - V[0:nz + 3, 1:ny + 4, 1:nx + 3]) * inv
    dwdy = (W[1:nz + 4, 1:ny + 4, 1:nx + 3] - W[1:nz + 4, 0:ny + 3, 1:nx + 3]) * inv

    def a_xy(q):
        return (q[:, :-1, :-1] + q[:, 1:, :-1] + q[:, :-1, 1:] + q[:, 1:, 1:]) * 0.25

    def a_xz(q):
        return (q[:-1, :, :-1] + q[1:, :, :-1] + q[:-1, :, 1:] + q[1:, :, 1:]) * 0.25

    def a_yz(q):
        return (q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]) * 0.25

    g = ((g11, a_xy(dudy), a_xz(dudz)), (a_xy(dvdx), g22, a_yz(dvdz)), (a_xz(dwdx), a_yz(dwdy), g33))
    return {"g": g, "sxy": dudy + dvdx, "sxz": dudz + dwdx, "syz": dvdz + dwdy}


def traceless(g):
    return (g[0], g[1], (g[2][0], g[2][1], -(g[0][0] + g[1][1])))


def eddy_viscosity3d(g, ops):
    """Smagorinsky / WALE / Vreman from g[i][j] = du_i/dx_j (core.py:456-488)."""
    delta2 = ops.h * ops.h
    c = ops.les_c
    s = [[0.5 * (g[i][j] + g[j][i]) for j in range(3)] for i in range(3)]
    ss = sum(s[i][j] * s[i][j] for i in range(3) for j in range(3))
    if ops.turbulence == "smagorinsky":
        return (c * c * delta2) * torch.sqrt(2.0 * ss + SQRT_EPS)
    if ops.turbulence == "wale":
        g2 = [[sum(g[i][k] * g[k][j] for k in range(3)) for j in range(3)] for i in range(3)]
        tr3 = (g2[0][0] + g2[1][1] + g2[2][2]) * (1.0 / 3.0)
        sd = 0.0
        for i in range(3):
            for j in range(3):
                sdij = 0.5 * (g2[i][j] + g2[j][i])
                if i == j:
                    sdij = sdij - tr3
                sd = sd + sdij * sdij
        num = sd.clamp_min(0.0) ** 1.5
        den = ss.clamp_min(0.0) ** 2.5 + sd.clamp_min(0.0) ** 1.25 + 1e-30
        return (c * c * delta2) * num / den
    if ops.turbulence == "vreman":
        b = [[sum(g[i][m] * g[j][m] for m in range(3)) for j in range(3)] for i in range(3)]
        bb = (b[0][0] * b[1][1] - b[0][1] * b[0][1] + b[0][0] * b[2][2] - b[0][2] * b[0][2]
              + b[1][1] * b[2][2] - b[1][2] * b[1][2]).clamp_min(0.0)
        aa = sum(g[i][j] * g[i][j] for i in range(3) for j in range(3))
        return (c * delta2) * torch.sqrt(bb / (aa + 1e-30) + SQRT_EPS)
    raise ValueError(ops.turbulence)


def learned_nut(nut, field, trust, ops):
    """The learned closure's eddy viscosity where it lies in [0, trust], the model's elsewhere
    (no reference counterpart). ``nut`` covers cells -1..n per axis; ``field`` (nz, ny, nx)
    wraps onto the ghost ring of periodic axes; non-periodic ghost cells keep the model."""
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    idx = []
    for n, per in zip((nz, ny, nx), reversed(ops.periodic_axes)):
        idx.append(torch.arange(-1, n + 1) % n if per else torch.arange(0, n))
    sub = field[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
    sl = tuple(slice(None) if per else slice(1, -1) for per in reversed(ops.periodic_axes))
    ok = (sub >= 0) & (sub <= trust)
    out = nut.clone()
    out[sl] = torch.where(ok, sub, nut[sl])
    return out


def sgs_divergence3d(U, V, W, geom, ops, nut_field=None, nut_trust=None):
    """div(2 nu_t S) at the faces, max nu_t (core.py:491-534); with ``nut_field`` the learned
    closure's values replace the model's where trusted (before the solid mask)."""
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    inv = 1.0 / ops.h
    gr = velocity_gradient(U, V, W, ops.h)
    g = gr["g"]
    nut = eddy_viscosity3d(traceless(g), ops)
    if nut_field is not None:
        nut = learned_nut(nut, nut_field, nut_trust, ops)
    if not geom.empty:
        k0, k1, j0, j1, i0, i1 = geom.window
        win = nut[k0 + 1:k1 + 1, j0 + 1:j1 + 1, i0 + 1:i1 + 1] * (1.0 - geom.chi_c)
        nut = nut.clone()
        nut[k0 + 1:k1 + 1, j0 + 1:j1 + 1, i0 + 1:i1 + 1] = win
    i = slice(1, -1)
    n_xy = nut[i]
    n_xy = (n_xy[:, :-1, :-1] + n_xy[:, 1:, :-1] + n_xy[:, :-1, 1:] + n_xy[:, 1:, 1:]) * 0.25
    n_xz = nut[:, i]
    n_xz = (n_xz[:-1, :, :-1] + n_xz[1:, :, :-1] + n_xz[:-1, :, 1:] + n_xz[1:, :, 1:]) * 0.25
    n_yz = nut[:, :, i]
    n_yz = (n_yz[:-1, :-1] + n_yz[1:, :-1] + n_yz[:-1, 1:] + n_yz[1:, 1:]) * 0.25
    t12 = n_xy * gr["sxy"][1:nz + 1, 1:ny + 2, 1:nx + 2]
    t13 = n_xz * gr["sxz"][1:nz + 2, 1:ny + 1, 1:nx + 2]
    t23 = n_yz * gr["syz"][1:nz + 2, 1:ny + 2, 1:nx + 1]
    t11 = 2.0 * nut[i, i] * g[0][0][i, i]
    t22 = 2.0 * nut[i, :, i] * g[1][1][i, :, i]
    t33 = 2.0 * nut[:, i, i] * g[2][2][:, i, i]
    fu = (t11[..., 1:] - t11[..., :-1] + t12[:, 1:] - t12[:, :-1] + t13[1:] - t13[:-1]) * inv
    fv = (t12[..., 1:] - t12[..., :-1] + t22[:, 1:] - t22[:, :-1] + t23[1:] - t23[:-1]) * inv
    fw = (t13[..., 1:] - t13[..., :-1] + t23[:, 1:] - t23[:, :-1] + t33[1:] - t33[:-1]) * inv
    return fu, fv, fw, nut.amax()


def boundary_tendencies(u, lu, lv, lw, ops):
    px, py, pz = ops.periodic_axes
    if not px:
        out = (u[..., -2:-1] - u[..., -1:]) * (1.0 / ops.h)
        if ops.closed:
            out = torch.zeros_like(out)
        lu = torch.cat([torch.zeros_like(out), lu[..., 1:-1], out], dim=2)
    if not py:
        z = torch.zeros_like(lv[:, :1])
        lv = torch.cat([z, lv[:, 1:-1], z], dim=1)
    if not pz:
        z = torch.zeros_like(lw[:1])
        lw = torch.cat([z, lw[1:-1], z], dim=0)
    return lu, lv, lw


def momentum_rhs3d(u, v, w, geom, ops, nut_field=None, nut_trust=None):
    U, V, W = pad_velocity(u, v, w, ops)
    au, av, aw = advection3(U, V, W, ops)
    nu_h2 = ops.nu * (1.0 / ops.h ** 2)
    lu = torch.addcmul(au, laplacian_padded(U), nu_h2)
    lv = torch.addcmul(av, laplacian_padded(V), nu_h2)
    lw = torch.addcmul(aw, laplacian_padded(W), nu_h2)
    nut_max = None
    if ops.turbulence != "none":
        fu, fv, fw, nut_max = sgs_divergence3d(U, V, W, geom, ops, nut_field, nut_trust)
        lu, lv, lw = lu + fu, lv + fv, lw + fw
    lu, lv, lw = boundary_tendencies(u, lu, lv, lw, ops)
    return lu, lv, lw, nut_max


# ---------------------------------------------------------------------------
# projection, penalisation, dt, forces (core.py:577-678)
# ---------------------------------------------------------------------------

def divergence3d(u, v, w, h):
    return ((u[..., 1:] - u[..., :-1]) + (v[:, 1:] - v[:, :-1]) + (w[1:] - w[:-1])) * (1.0 / h)


def pressure_gradient(psi, ops):
    inv = 1.0 / ops.h
    px, py, pz = ops.periodic_axes
    if px:
        gx = (psi - torch.roll(psi, 1, dims=2)) * inv
        gx = torch.cat([gx, gx[..., :1]], dim=2)
    else:
        gx = F.pad((psi[..., 1:] - psi[..., :-1]) * inv, (1, 1))
    if py:
        gy = (psi - torch.roll(psi, 1, dims=1)) * inv
        gy = torch.cat([gy, gy[:, :1]], dim=1)
    else:
        gy = F.pad((psi[:, 1:] - psi[:, :-1]) * inv, (0, 0, 1, 1))
    if pz:
        gz = (psi - torch.roll(psi, 1, dims=0)) * inv
        gz = torch.cat([gz, gz[:1]], dim=0)
    else:
        gz = F.pad((psi[1:] - psi[:-1]) * inv, (0, 0, 0, 0, 1, 1))
    return gx, gy, gz


def project3d(u, v, w, ops):
    psi = poisson(ops).solve(divergence3d(u, v, w, ops.h))
    gx, gy, gz = pressure_gradient(psi, ops)
    return u - gx, v - gy, w - gz, psi


def penalise3d(u, v, w, geom, ops, bdt):
    k0, k1, j0, j1, i0, i1 = geom.window
    k = bdt / ops.eta
    u, v, w = u.clone(), v.clone(), w.clone()
    uw = u[k0:k1, j0:j1, i0:i1 + 1]
    vw = v[k0:k1, j0:j1 + 1, i0:i1]
    ww = w[k0:k1 + 1, j0:j1, i0:i1]
    un = (uw + k * geom.a_u) / (1.0 + k * geom.chi_u)
    vn = (vw + k * geom.a_v) / (1.0 + k * geom.chi_v)
    wn = (ww + k * geom.a_w) / (1.0 + k * geom.chi_w)
    du, dv, dw = un - uw, vn - vw, wn - ww
    u[k0:k1, j0:j1, i0:i1 + 1] = un
    v[k0:k1, j0:j1 + 1, i0:i1] = vn
    w[k0:k1 + 1, j0:j1, i0:i1] = wn
    return u, v, w, du, dv, dw


def cfl_dt3d(u, v, w, ops, nut_max=None):
    umax = (u.abs().amax() + v.abs().amax() + w.abs().amax()).clamp_min(1e-6)
    dt = ops.cfl * ops.h / umax
    nu = ops.nu if nut_max is None else ops.nu + nut_max
    return torch.minimum(dt, VISCOUS_DT * ops.h * ops.h / nu)


def force_coefficients(acc, geom, ops, dt):
    """[CD, CS, CL, C_roll, Cm, Cn] per body (core.py:653-678)."""
    acc_u, acc_v, acc_w = acc
    scale = -(ops.h ** 3) / dt
    ru = torch.einsum("bkzyx,zyx->bk", geom.k_u, acc_u) * scale
    rv = torch.einsum("bkzyx,zyx->bk", geom.k_v, acc_v) * scale
    rw = torch.einsum("bkzyx,zyx->bk", geom.k_w, acc_w) * scale
    fx, fy, fz = ru[:, 0], rv[:, 0], rw[:, 0]
    mx, my, mz = rv[:, 1] + rw[:, 1], ru[:, 1] + rw[:, 2], ru[:, 2] + rv[:, 2]
    if ops.symmetry:
        dy = ops.moment_ref[1] - ops.y0
        fx, fz, my = 2.0 * fx, 2.0 * fz, 2.0 * my
        fy = torch.zeros_like(fy)
        mx = -dy * fz
        mz = dy * fx
    q_a = 0.5 * ops.ref_area
    return torch.stack([fx / q_a, fy / q_a, fz / q_a, mx / (q_a * ops.span_ref), my / (q_a * ops.length_ref),
                        mz / (q_a * ops.span_ref)], dim=1)


def step3d_full(state: State3D, geom, ops, dt=None, nut_field=None, nut_trust=None):
    """One SSP-RK3 step without placed physics (core.py:765-883); ``nut_field`` / ``nut_trust``:
    the learned closure's eddy viscosity, held over the step (no reference counterpart)."""
    u0, v0, w0 = state.u, state.v, state.w
    up, vp, wp = u0, v0, w0
    acc = None
    psi_w = None
    px = ops.periodic_axes[0]
    any_periodic = any(ops.periodic_axes)
    for s, ((a, b), wgt) in enumerate(zip(SSP_RK3, FORCE_WEIGHTS)):
        lu, lv, lw, nut_max = momentum_rhs3d(up, vp, wp, geom, ops, nut_field, nut_trust)
        if s == 0:
            if dt is None:
                dt = ops.extras["dt_fixed"] if ops.dt_fixed is not None else cfl_dt3d(u0, v0, w0, ops, nut_max)
            if not torch.is_tensor(dt):
                dt = torch.tensor(float(dt), dtype=ops.dtype)
        bdt = dt * b
        if a == 0.0:
            uh, vh, wh = torch.addcmul(up, lu, bdt), torch.addcmul(vp, lv, bdt), torch.addcmul(wp, lw, bdt)
        else:
            uh = torch.addcmul(torch.lerp(up, u0, a), lu, bdt)
            vh = torch.addcmul(torch.lerp(vp, v0, a), lv, bdt)
            wh = torch.addcmul(torch.lerp(wp, w0, a), lw, bdt)
        if not px:
            uh = inlet_outlet(uh, ops)
        if not geom.empty:
            uh, vh, wh, du, dv, dw = penalise3d(uh, vh, wh, geom, ops, bdt)
            acc = (du * wgt, dv * wgt, dw * wgt) if acc is None else (
                acc[0] + du * wgt, acc[1] + dv * wgt, acc[2] + dw * wgt)
            if any_periodic:
                uh, vh, wh = sync_periodic(uh, vh, wh, ops)
        up, vp, wp, psi = project3d(uh, vh, wh, ops)
        psi_w = psi * wgt if psi_w is None else torch.add(psi_w, psi, alpha=wgt)
    p = psi_w * (1.0 / dt)
    if geom.empty:
        coeffs = torch.zeros(geom.nbody, 6, dtype=ops.dtype)
    else:
        coeffs = force_coefficients(acc, geom, ops, dt)
    return State3D(up, vp, wp, p, state.t + dt, dt * torch.ones_like(state.t)), coeffs


def noisy_start3d(ops, seed, amplitude=0.01):
    """Inflow + seeded Gaussian noise on every face, boundary values, projection."""
    import numpy as np
    rng = np.random.default_rng(seed)
    nz, ny, nx = ops.nz, ops.ny, ops.nx
    nu_ = rng.standard_normal((nz, ny, nx + 1))
    nv_ = rng.standard_normal((nz, ny + 1, nx))
    nw_ = rng.standard_normal((nz + 1, ny, nx))
    base = 0.0 if (all(ops.periodic_axes) or ops.closed) else 1.0
    u = torch.as_tensor(base + amplitude * nu_, dtype=ops.dtype)
    v = torch.as_tensor(amplitude * nv_, dtype=ops.dtype)
    w = torch.as_tensor(amplitude * nw_, dtype=ops.dtype)
    u, v, w = apply_boundary_values(u, v, w, ops)
    u, v, w, _ = project3d(u, v, w, ops)
    return u, v, w


# ---------------------------------------------------------------------------
# smoke (ns3d/dye.py:232-305)
# ---------------------------------------------------------------------------

DYE_FLOOR = 1e-8                                          # ns3d/dye.py:63


def pad_dye(c, ops):
    """Two ghost layers per face (dye.py:232-263): x clean inflow / zero-gradient outlet,
    y and z mirrored, periodic axes wrapped; edges and corners unused (left at zero)."""
    nz, ny, nx = c.shape
    P = c.new_zeros(nz + 4, ny + 4, nx + 4)
    P[2:-2, 2:-2, 2:-2] = c
    px, py, pz = ops.periodic_axes
    m = slice(2, -2)
    if px:
        P[m, m, 0:2] = c[..., nx - 2:]
        P[m, m, nx + 2:] = c[..., :2]
    else:
        P[m, m, 0:2] = 0.0
        P[m, m, nx + 2:] = c[..., nx - 1:].expand(nz, ny, 2)
    if py:
        P[m, 0:2, m] = c[:, ny - 2:]
        P[m, ny + 2:, m] = c[:, :2]
    else:
        P[m, 0:2, m] = c[:, :2].flip(1)
        P[m, ny + 2:, m] = c[:, ny - 2:].flip(1)
    if pz:
        P[0:2, m, m] = c[nz - 2:]
        P[nz + 2:, m, m] = c[:2]
    else:
        P[0:2, m, m] = c[:2].flip(0)
        P[nz + 2:, m, m] = c[nz - 2:].flip(0)
    return P


def _van_leer(p, q):
    pq = p * q
    pos = pq > 0
    return torch.where(pos, 2.0 * pq / torch.where(pos, p + q, torch.ones_like(p)), torch.zeros_like(p))


def dye_face_flux(a, b, c, d, vel):
    """Upwind MUSCL flux through the face between b and c (dye.py:274-278)."""
    up = b + 0.5 * _van_leer(b - a, c - b)
    dn = c - 0.5 * _van_leer(c - b, d - c)
    return vel * torch.where(vel > 0, up, dn)


def advect_dye3d(c, u, v, w, dt, ops):
    """Forward Euler with MUSCL / van Leer fluxes, clipped at 1, floored (dye.py:281-297)."""
    nz, ny, nx = c.shape
    P = pad_dye(c, ops)
    m = slice(2, -2)
    fx = dye_face_flux(P[m, m, 0:nx + 1], P[m, m, 1:nx + 2], P[m, m, 2:nx + 3], P[m, m, 3:nx + 4], u)
    fy = dye_face_flux(P[m, 0:ny + 1, m], P[m, 1:ny + 2, m], P[m, 2:ny + 3, m], P[m, 3:ny + 4, m], v)
    fz = dye_face_flux(P[0:nz + 1, m, m], P[1:nz + 2, m, m], P[2:nz + 3, m, m], P[3:nz + 4, m, m], w)
    k = dt * (1.0 / ops.h)
    r = c - k * ((fx[..., 1:] - fx[..., :-1]) + (fy[:, 1:] - fy[:, :-1]) + (fz[1:] - fz[:-1]))
    r = r.clamp(max=1.0)
    return torch.where(r < DYE_FLOOR, torch.zeros_like(r), r)


def apply_solids_and_sources(c, t, geom, window, weight, pulse):
    """In place: zero where the solid fraction exceeds 1/2, then c += s (1 - c) in the nozzles
    with s = weight x [t mod period < t_on] (dye.py:300-312)."""
    if geom is not None and not geom.empty:
        k0, k1, j0, j1, i0, i1 = geom.window
        c[k0:k1, j0:j1, i0:i1].masked_fill_(geom.chi_c > 0.5, 0.0)
    k0, k1, j0, j1, i0, i1 = window
    if k1 > k0 and j1 > j0 and i1 > i0:
        win = c[k0:k1, j0:j1, i0:i1]
        on = (torch.remainder(t, pulse[1]) < pulse[0]).to(c.dtype)
        win.add_(weight * on * (1.0 - win))
