"""CPU checks of the learned closure's oracle rule and of its data functions against what
the mathematics fixes.

* ``oracle.ns3d.learned_nut`` inside a step: a field that is untrusted everywhere gives the
  WALE step bit for bit; a field trusted everywhere at a constant c gives the laminar step
  with viscosity nu + c (on a discretely divergence-free periodic field the stress
  divergence of a constant eddy viscosity is c times the Laplacian), to round-off.
* ``closure.data``: the spectrum of one Fourier mode sits in one shell with E = A^2 / 4; the
  face filter is the block mean (exact on linear fields) and keeps a divergence-free field
  divergence-free; ``sgs_target`` recovers the viscosity of a planted eddy-viscosity stress
  and clips a backscatter stress to 0; ``sgs_stress`` vanishes on fields constant over each
  filter box; the seeded initial field has u_rms = 1, no divergence and the target spectrum
  shape to sampling error."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import ns3d as O


def _cfg(turbulence="wale", re=300.0, n=(16, 12, 10)):
    from aerolume_amd.ns3d.config import SolverConfig3D
    nx, ny, nz = n
    return SolverConfig3D(nx=nx, ny=ny, nz=nz, h=2 * math.pi / nx, x0=0, y0=0, z0=0, re=re, periodic=True,
                          turbulence=turbulence, dtype="float64", device="cpu", dye=False, perturb=False)


def _ops(cfg):
    from aerolume_amd.ns3d import setup as su
    return su.make_operators3d(cfg, device="cpu")


def _start(ops, seed=21):
    u, v, w = O.noisy_start3d(ops, seed, amplitude=0.5)
    z = torch.zeros((), dtype=torch.float64)
    return O.State3D(u, v, w, torch.zeros(ops.nz, ops.ny, ops.nx, dtype=torch.float64), z, z)


def _geom():
    from aerolume_amd.ns3d import setup as su
    return su.empty_geometry(0)


def test_untrusted_learned_field_is_the_wale_step():
    ops = _ops(_cfg())
    st = _start(ops)
    ref, _ = O.step3d_full(st, _geom(), ops)
    for bad in (-1.0, 7.0):           # below 0 or above the trust bound everywhere
        field = torch.full((ops.nz, ops.ny, ops.nx), bad, dtype=torch.float64)
        got, _ = O.step3d_full(st, _geom(), ops, nut_field=field, nut_trust=1.0)
        for k in ("u", "v", "w", "p", "dt"):
            assert torch.equal(getattr(got, k), getattr(ref, k)), (bad, k)


def test_trusted_constant_field_is_a_viscosity_shift():
    c, re = 0.004, 300.0
    ops_w = _ops(_cfg("wale", re))
    st = _start(ops_w)
    field = torch.full((ops_w.nz, ops_w.ny, ops_w.nx), c, dtype=torch.float64)
    got, _ = O.step3d_full(st, _geom(), ops_w, nut_field=field, nut_trust=1.0)
    ops_l = _ops(_cfg("none", 1.0 / (1.0 / re + c)))
    ref, _ = O.step3d_full(st, _geom(), ops_l)
    assert abs(float(got.dt) - float(ref.dt)) < 1e-15
    for k in ("u", "v", "w"):
        assert float((getattr(got, k) - getattr(ref, k)).abs().max()) < 1e-12, k
    assert float((got.p - ref.p).abs().max()) < 1e-9
    # a planted mistake: the WALE value where the field is trusted moves the result
    wale, _ = O.step3d_full(st, _geom(), ops_w)
    assert float((wale.u - ref.u).abs().max()) > 1e-6


def _stag(n, fu=None, fv=None, fw=None):
    z = torch.zeros
    u = z(n, n, n + 1, dtype=torch.float64) if fu is None else fu
    v = z(n, n + 1, n, dtype=torch.float64) if fv is None else fv
    w = z(n + 1, n, n, dtype=torch.float64) if fw is None else fw
    return u, v, w


def test_spectrum_of_one_mode():
    from aerolume_amd.closure import data as D
    n, m, a = 16, 3, 0.7
    zc = (torch.arange(n, dtype=torch.float64) + 0.5) * (2 * math.pi / n)
    vz = a * torch.sin(m * zc)
    v = vz[:, None, None].expand(n, n + 1, n).clone()      # a function of z only: centring is exact
    k, e = D.kinetic_energy_spectrum(*_stag(n, fv=v))
    assert abs(e[m] - a * a / 4) < 1e-14
    assert float(np.abs(np.delete(e, m)).max()) < 1e-28
    assert abs(e.sum() - float((vz * vz).mean()) / 2) < 1e-14


def test_face_filter_block_mean_and_divergence():
    from aerolume_amd.closure import data as D
    n, r = 16, 4
    kk, jj, ii = torch.meshgrid(*(torch.arange(s, dtype=torch.float64) for s in (n, n, n + 1)), indexing="ij")
    u = 0.5 + 0.25 * jj - 0.125 * kk + 0.0 * ii
    uc, vc, wc = D.face_filter(*_stag(n, fu=u), r)
    m = n // r
    K, J = torch.meshgrid(torch.arange(m, dtype=torch.float64), torch.arange(m, dtype=torch.float64), indexing="ij")
    want = 0.5 + 0.25 * (J * r + (r - 1) / 2) - 0.125 * (K * r + (r - 1) / 2)
    assert float((uc[..., 0] - want).abs().max()) < 1e-14 and uc.shape == (m, m, m + 1)
    # a divergence-free periodic field stays divergence-free on the coarse grid
    ops = _ops(_cfg(n=(n, n, n)))
    st = _start(ops, seed=4)
    assert float(O.divergence3d(st.u, st.v, st.w, ops.h).abs().max()) < 1e-12
    uc, vc, wc = D.face_filter(st.u, st.v, st.w, r)
    assert float(O.divergence3d(uc, vc, wc, ops.h * r).abs().max()) < 1e-12


def test_sgs_target_recovers_planted_viscosity():
    from aerolume_amd.closure import data as D
    rng = np.random.default_rng(2)
    g = torch.as_tensor(rng.standard_normal((9, 3, 4, 5)))
    G = g.view(3, 3, 3, 4, 5)
    s = 0.5 * (G + G.transpose(0, 1))
    nu0, iso = 0.03, 0.2
    tau_full = -2.0 * nu0 * s + iso * torch.eye(3, dtype=torch.float64)[:, :, None, None, None]
    tau = torch.stack([tau_full[0, 0], tau_full[1, 1], tau_full[2, 2], tau_full[0, 1], tau_full[0, 2], tau_full[1, 2]])
    got = D.sgs_target(tau, g, eps=0.0)
    # the deviatoric part of -2 nu0 S is -2 nu0 (S - tr S / 3): its contraction with S over
    # 2 S : S is nu0 (1 - (tr S)^2 / (3 S : S))
    ss = (s * s).sum(dim=(0, 1))
    trs = s[0, 0] + s[1, 1] + s[2, 2]
    want = nu0 * (1.0 - trs * trs / (3.0 * ss))
    assert float((got - want).abs().max()) < 1e-15
    assert float(D.sgs_target(-tau, g, eps=0.0).abs().max()) == 0.0      # backscatter clipped


def test_sgs_stress_vanishes_on_box_constant_fields():
    from aerolume_amd.closure import data as D
    n, r = 8, 4
    rng = np.random.default_rng(3)
    blk = torch.as_tensor(rng.standard_normal((n // r,) * 3))
    # v and w constant along x, y, z within each filter box except along their own normal,
    # where a face value shared by two boxes would mix them: take them constant along the
    # whole normal axis, and u uniform
    col = blk.repeat_interleave(r, 0).repeat_interleave(r, 1).repeat_interleave(r, 2)
    v = torch.cat([col[:, :1]] * (n + 1), dim=1).contiguous()          # a function of (z, x) only
    w = torch.cat([col[:1]] * (n + 1), dim=0).contiguous()             # a function of (y, x) only
    u = torch.full((n, n, n + 1), 1.5, dtype=torch.float64)
    tau = D.sgs_stress(u, v, w, r)
    assert float(tau.abs().max()) < 1e-14
    # and a field varying inside the boxes does carry a stress (the check is not vacuous)
    tau2 = D.sgs_stress(u, torch.roll(v, 1, dims=2), w, r)
    assert float(tau2.abs().max()) > 1e-3


def test_seeded_field_statistics():
    from aerolume_amd.closure import data as D
    n = 32
    u = D.spectral_velocity(n, 505)
    assert abs(math.sqrt(float((u * u).sum(0).mean()) / 3.0) - 1.0) < 1e-12
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    f = [np.fft.fftn(c) for c in u]
    div = kx * f[0] + ky * f[1] + kz * f[2]
    # solenoidal below the Nyquist planes (there a real field's mode pairs with itself under
    # k -> -k, and the derivative is not defined)
    inner = (np.abs(kx) < n // 2) & (np.abs(ky) < n // 2) & (np.abs(kz) < n // 2)
    assert float(np.abs(div[inner]).max()) < 1e-12 * float(np.abs(f[0]).max())
    kk = np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)).astype(int)
    e = np.zeros(kk.max() + 1)
    for c in f:
        np.add.at(e, kk.ravel(), 0.5 * (np.abs(c / c.size) ** 2).ravel())
    ks = np.arange(2, 9)
    target = ks ** 4.0 * np.exp(-2.0 * (ks / 4.0) ** 2)
    ratio = e[ks] / target
    # shell sums of about 4 pi k^2 random modes: within 35 % of the fitted level
    assert np.abs(ratio / ratio.mean() - 1.0).max() < 0.35


def test_wale_baseline_is_the_solver_closure():
    """``closure.data.wale_nut`` (the a-priori baseline) is the oracle's WALE eddy viscosity
    divided by its constant factor, on random gradients."""
    from aerolume_amd.closure import data as D
    ops = _ops(_cfg("wale"))
    g = torch.as_tensor(np.random.default_rng(8).standard_normal((9, 3, 4, 5)))
    ref = O.eddy_viscosity3d(O.traceless([[g[3 * i + j] for j in range(3)] for i in range(3)]), ops)
    ratio = ref / D.wale_nut(g)
    assert float((ratio / ratio.mean() - 1.0).abs().max()) < 1e-12
