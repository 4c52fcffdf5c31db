"""Training data of the learned closure: decaying turbulence in a triply periodic box solved
on a fine grid by this package's 3D step, box-filtered to a coarse grid.

* ``spectral_velocity``: a seeded solenoidal field with E(k) proportional to
  k^4 exp(-2 (k / k0)^2), scaled to u_rms = 1, on the staggered grid (projected exactly).
* ``face_filter``: restriction of a staggered field to a grid r times coarser: each coarse
  face value is the mean of the r x r fine faces it covers (flux conserving, so a discretely
  divergence-free fine field stays divergence-free on the coarse grid).
* ``sgs_target``: the eddy viscosity that matches the resolved-to-subgrid energy transfer of
  the filtered field, nu = max(0, -tau^d : S / (2 S : S)), with tau = box(u u) - box(u) box(u).
* ``kinetic_energy_spectrum``: shell-summed E(k) of the cell-centred velocity (host FFT).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from ..ns3d import solver as S3
from ..ns3d.config import SolverConfig3D


def box_config(n: int, re: float = 1000.0, turbulence: str = "none", dtype: str = "float32", cfl: float = 0.5,
               advection: str = "upwind3"):
    """The triply periodic [0, 2 pi)^3 box at n^3 cells, nu = 1 / re (U = u_rms = 1)."""
    return SolverConfig3D(nx=n, ny=n, nz=n, h=2.0 * math.pi / n, x0=0.0, y0=0.0, z0=0.0, re=re, periodic=True,
                          turbulence=turbulence, advection=advection, cfl=cfl, dtype=dtype, device="cuda", dye=False,
                          perturb=False)


def spectral_velocity(n: int, seed: int, k0: float = 4.0):
    """Cell-centred solenoidal velocity (3, n, n, n) float64 with the target spectrum, u_rms = 1."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, kr, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    e = kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    amp = np.where(kk > 0, np.sqrt(e / (4.0 * math.pi * np.maximum(kk, 1e-30) ** 2)), 0.0)
    c = [(rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)) * amp for _ in range(3)]
    # solenoidal projection in Fourier space
    k2 = np.maximum(kk * kk, 1e-30)
    div = (kx * c[0] + ky * c[1] + kz * c[2]) / k2
    c = [c[0] - kx * div, c[1] - ky * div, c[2] - kz * div]
    u = np.stack([np.fft.irfftn(ci, s=(n, n, n), axes=(0, 1, 2)) for ci in c])   # (comp, z, y, x) with comp x, y, z
    urms = math.sqrt(float((u * u).sum(0).mean()) / 3.0)
    return u / urms


def to_staggered(uc: np.ndarray, ops):
    """Cell-centred (3, n, n, n) -> staggered (u, v, w) on the device by face averages, with the
    periodic duplicate faces, then the exact projection."""
    ux, uy, uz = (torch.as_tensor(a, dtype=ops.dtype, device=ops.device) for a in uc)
    u = 0.5 * (ux + torch.roll(ux, 1, dims=2))
    v = 0.5 * (uy + torch.roll(uy, 1, dims=1))
    w = 0.5 * (uz + torch.roll(uz, 1, dims=0))
    u = torch.cat([u, u[..., :1]], dim=2).contiguous()
    v = torch.cat([v, v[:, :1]], dim=1).contiguous()
    w = torch.cat([w, w[:1]], dim=0).contiguous()
    u, v, w, _ = S3.project3d(u, v, w, ops)
    return u, v, w


def face_filter(u, v, w, r: int):
    """Staggered periodic fields (with duplicate last faces) -> the r-times coarser staggered
    fields: mean of the r x r fine faces on each coarse face."""
    nz, ny, nxp = u.shape
    nx = nxp - 1
    Z, Y, X = nz // r, ny // r, nx // r
    uc = u[..., :nx:r].reshape(Z, r, Y, r, X).mean(dim=(1, 3))
    vc = v[:, :ny:r, :].reshape(Z, r, Y, X, r).mean(dim=(1, 4))
    wc = w[:nz:r].reshape(Z, Y, r, X, r).mean(dim=(2, 4))
    uc = torch.cat([uc, uc[..., :1]], dim=2).contiguous()
    vc = torch.cat([vc, vc[:, :1]], dim=1).contiguous()
    wc = torch.cat([wc, wc[:1]], dim=0).contiguous()
    return uc, vc, wc


def _box(q, r: int):
    n = q.shape[-1]
    m = n // r
    return q.reshape(*q.shape[:-3], m, r, m, r, m, r).mean(dim=(-5, -3, -1))


def sgs_stress(u, v, w, r: int):
    """tau_ij = box(u_i u_j) - box(u_i) box(u_j) of the fine cell-centred velocity, on the coarse
    cells: (6, Z, Y, X) in the order xx, yy, zz, xy, xz, yz."""
    uc = S3.centre_u(u)
    vc = S3.centre_v(v)
    wc = S3.centre_w(w)
    comps = (uc, vc, wc)
    bar = [_box(c, r) for c in comps]
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    return torch.stack([_box(comps[a] * comps[b], r) - bar[a] * bar[b] for a, b in pairs])


def sgs_target(tau, g, eps: float = 1e-12):
    """nu = max(0, -tau^d : S / (2 S : S)) from the SGS stress (6, ...) and the coarse
    cell-centred gradient g (9, ...) = du_i/dx_j."""
    G = g.view(3, 3, *g.shape[1:])
    s = 0.5 * (G + G.transpose(0, 1))
    tr = (tau[0] + tau[1] + tau[2]) / 3.0
    t = torch.stack([torch.stack([tau[0] - tr, tau[3], tau[4]]), torch.stack([tau[3], tau[1] - tr, tau[5]]),
                     torch.stack([tau[4], tau[5], tau[2] - tr])])
    num = -(t * s).sum(dim=(0, 1))
    den = 2.0 * (s * s).sum(dim=(0, 1)) + eps
    return torch.clamp(num / den, min=0.0)


def wale_nut(g):
    """WALE eddy viscosity up to its constant factor (Cw h)^2 from a cell-centred gradient
    (9, ...) = du_i/dx_j: sd^1.5 / (S:S^2.5 + sd^1.25), sd = |S^d|^2 of the traceless square
    of the gradient (the solver's closure on the same gradient, made traceless as the solver
    does, by du_3/dx_3 = -(du_1/dx_1 + du_2/dx_2); the factor drops out of a correlation)."""
    G = g.view(3, 3, *g.shape[1:]).clone()
    G[2, 2] = -(G[0, 0] + G[1, 1])
    eye = torch.eye(3, dtype=g.dtype, device=g.device).view(3, 3, *([1] * (g.dim() - 1)))
    s = 0.5 * (G + G.transpose(0, 1))
    g2 = torch.einsum("ik...,kj...->ij...", G, G)
    sd = 0.5 * (g2 + g2.transpose(0, 1)) - ((g2[0, 0] + g2[1, 1] + g2[2, 2]) / 3.0) * eye
    ss = (s * s).sum(dim=(0, 1)).clamp_min(0.0)
    sdd = (sd * sd).sum(dim=(0, 1)).clamp_min(0.0)
    return sdd ** 1.5 / (ss ** 2.5 + sdd ** 1.25 + 1e-30)


def kinetic_energy_spectrum(u, v, w) -> tuple[np.ndarray, np.ndarray]:
    """(k, E(k)) of the cell-centred velocity of a periodic staggered field, E summed over
    integer shells |k| in [k - 1/2, k + 1/2), normalised so that sum E = mean(|u|^2) / 2."""
    comps = [S3.centre_u(u).double().cpu().numpy(), S3.centre_v(v).double().cpu().numpy(),
             S3.centre_w(w).double().cpu().numpy()]
    n = comps[0].shape[0]
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    shell = np.rint(kk).astype(np.int64)
    e = np.zeros(shell.max() + 1)
    for c in comps:
        f = np.fft.fftn(c) / c.size
        np.add.at(e, shell.ravel(), 0.5 * (np.abs(f) ** 2).ravel())
    return np.arange(len(e)), e
