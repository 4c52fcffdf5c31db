"""CPU checks of the 2D oracle and set-up code against the reference's golden vectors and
against what the mathematics fixes (closed forms, invariants, brute force)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import ns2d as O
from tests import cases


def _ops(cfg):
    from aerolume_amd.ns2d import setup as su
    return su.make_operators(cfg, device="cpu")


def _geom(cfg, ops, live=False):
    from aerolume_amd.ns2d import setup as su
    w = su.live_window(cfg.bodies, ops) if live else None
    return su.build_geometry(cfg.bodies, ops, window=w)


def _state(u, v, ops, theta=None, dye=None):
    z = torch.zeros((), dtype=ops.dtype)
    t = lambda a: None if a is None else torch.as_tensor(a, dtype=ops.dtype)  # noqa: E731
    return O.State(t(u), t(v), torch.zeros(ops.ny, ops.nx, dtype=ops.dtype), t(dye), z.clone(), z.clone(), t(theta))


# ----------------------------------------------------------------- Poisson
@pytest.mark.parametrize("nx,ny", [(16, 8), (15, 9), (12, 10)])
@pytest.mark.parametrize("periodic", [False, True])
def test_poisson_exact_residual(nx, ny, periodic):
    """L psi = f - mean(f) to 1e-10 relative, psi has zero mean (odd and even sizes)."""
    f = torch.as_tensor(np.random.default_rng(1).standard_normal((ny, nx)))
    ps = O.PoissonSolver(nx, ny, 0.3, periodic=periodic, dtype=torch.float64)
    psi = ps.solve(f)
    r = ps.laplacian(psi) - (f - f.mean())
    assert float(r.abs().max()) < 1e-10 * float(f.abs().max())
    assert abs(float(psi.mean())) < 1e-12


def test_dct_matches_scipy_and_round_trips():
    """Makhoul DCT equals scipy's DCT-II (unnormalised / 2) and idct inverts it."""
    from scipy.fft import dct
    x = torch.as_tensor(np.random.default_rng(2).standard_normal((5, 12)))
    for dim in (0, 1):
        X = O.dct2(x, dim)
        ref = dct(x.numpy(), type=2, axis=dim) / 2.0
        assert np.abs(X.numpy() - ref).max() < 1e-12
        assert float((O.idct2(X, dim) - x).abs().max()) < 1e-12


def test_poisson_matches_dense_least_squares():
    """Brute force: the dense 5-point Neumann operator solved by least squares (min norm)."""
    nx, ny, h = 6, 5, 0.5
    n = nx * ny
    L = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            for dj, di in ((0, 1), (0, -1), (1, 0), (-1, 0)):
                jj, ii = j + dj, i + di
                if 0 <= jj < ny and 0 <= ii < nx:
                    L[k, jj * nx + ii] += 1 / h**2
                    L[k, k] -= 1 / h**2
    f = np.random.default_rng(5).standard_normal(n)
    f -= f.mean()
    ref = np.linalg.lstsq(L, f, rcond=None)[0]
    psi = O.PoissonSolver(nx, ny, h, periodic=False, dtype=torch.float64).solve(torch.as_tensor(f.reshape(ny, nx)))
    assert np.abs(psi.numpy().ravel() - (ref - ref.mean())).max() < 1e-10


# ----------------------------------------------------------------- projection / advection
def test_projection_of_noise_is_divergence_free():
    cfg = cases.trace_cfg("float64")
    ops = _ops(cfg)
    rng = np.random.default_rng(4)
    u = torch.as_tensor(rng.standard_normal((ops.ny, ops.nx + 1)))
    v = torch.as_tensor(rng.standard_normal((ops.ny + 1, ops.nx)))
    u, v = O.apply_boundary_values(u, v, ops)
    un, vn, _, _ = O.project(u, v, ops)
    assert float(O.divergence(un, vn, ops.h).abs().max()) * ops.h < 1e-10
    assert torch.equal(un[:, 0], u[:, 0]) and torch.equal(un[:, -1], u[:, -1])


def test_upwind3_reconstruction_exact_on_quadratics():
    """upwind3 face flux x 7/24 = (vsum / 2) phi_face with phi_face = (-a + 5b + 2c)/6 for
    flow from the left (mirror from the right): exact for quadratics given as cell averages
    (reference core.py:375-380)."""
    Q = lambda x: 0.7 * x**3 / 3 - 1.3 * x**2 / 2 + 0.2 * x  # noqa: E731  (antiderivative)
    q0 = 0.2
    a, b, c, d = (torch.tensor(Q(x + 1) - Q(x), dtype=torch.float64) for x in (-2.0, -1.0, 0.0, 1.0))
    vsum = torch.tensor(2.0, dtype=torch.float64)
    flux = O.face_flux(vsum, a, b, c, d, True) * (7.0 / 24.0)
    assert abs(float(flux) - 0.5 * float(vsum) * q0) < 1e-12
    flux_neg = O.face_flux(-vsum, a, b, c, d, True) * (7.0 / 24.0)
    assert abs(float(flux_neg) + 0.5 * float(vsum) * q0) < 1e-12
    # a dropped dissipation term would only be second order: it fails on this quadratic
    cen = O.face_flux(vsum, a, b, c, d, False) * 0.25
    assert abs(float(cen) - 0.5 * float(vsum) * q0) > 1e-3


def _taylor_green(n, advection, steps, re=100.0):
    from aerolume_amd.ns2d.config import SolverConfig
    L = 2 * math.pi
    cfg = SolverConfig(nx=n, ny=n, h=L / n, x0=0.0, y0=0.0, re=re, periodic=True, advection=advection,
                       dtype="float64", device="cpu", dye=False, perturb=False, cfl=0.3)
    ops = _ops(cfg)
    geom = _geom(cfg, ops)
    h = ops.h
    xu, yu = np.meshgrid(np.arange(n + 1) * h, (np.arange(n) + 0.5) * h)
    xv, yv = np.meshgrid((np.arange(n) + 0.5) * h, np.arange(n + 1) * h)
    u = np.sin(xu) * np.cos(yu)
    v = -np.cos(xv) * np.sin(yv)
    st = _state(u, v, ops)
    for _ in range(steps):
        st, _ = O.step_full(st, geom, ops)
    t = float(st.t)
    decay = math.exp(-2.0 * t / re)
    xu, yu = np.meshgrid(np.arange(n + 1) * h, (np.arange(n) + 0.5) * h)
    err = np.abs(st.u.numpy() - decay * np.sin(xu) * np.cos(yu)).max()
    return err, t


def test_taylor_green_decay_and_second_order():
    """Periodic Taylor-Green vortex decays as exp(-2 t / Re); central advection converges
    at second order (errors at matched time)."""
    e1, t1 = _taylor_green(16, "central", 20)
    e2, t2 = _taylor_green(32, "central", 40)
    assert e1 < 5e-3 and e2 < 2e-3
    assert abs(t1 - t2) / t1 < 0.05
    assert e1 / e2 > 3.0


def test_central_scheme_conserves_energy_inviscid():
    from aerolume_amd.ns2d.config import SolverConfig
    n = 32
    cfg = SolverConfig(nx=n, ny=n, h=2 * math.pi / n, x0=0, y0=0, re=1e12, periodic=True, advection="central",
                       dtype="float64", device="cpu", dye=False, perturb=False, cfl=0.2)
    ops = _ops(cfg)
    geom = _geom(cfg, ops)
    rng = np.random.default_rng(9)
    u = torch.as_tensor(rng.standard_normal((n, n + 1)))
    v = torch.as_tensor(rng.standard_normal((n + 1, n)))
    u, v = O.apply_boundary_values(u, v, ops)
    u, v, _, _ = O.project(u, v, ops)
    e0 = float((u[:, :-1] ** 2).sum() + (v[:-1] ** 2).sum())
    st = _state(u, v, ops)
    for _ in range(20):
        st, _ = O.step_full(st, geom, ops)
    e1 = float((st.u[:, :-1] ** 2).sum() + (st.v[:-1] ** 2).sum())
    assert abs(e1 - e0) / e0 < 1e-4


# ----------------------------------------------------------------- LES closed forms
def _les_ops(model):
    from types import SimpleNamespace
    return SimpleNamespace(h=0.1, les_c={"smagorinsky": 0.17, "wale": 0.5, "vreman": 0.07}[model], turbulence=model)


@pytest.mark.parametrize("model", ["smagorinsky", "wale", "vreman"])
def test_eddy_viscosity_closed_forms(model):
    """nu_t >= 0; pure shear gives 0 for WALE and Vreman, (Cs h)^2 |S| sqrt2 for Smagorinsky."""
    ops = _les_ops(model)
    g = torch.as_tensor(np.random.default_rng(0).standard_normal((4, 50)))
    nut = O.eddy_viscosity(g[0], g[1], g[2], g[3], ops)
    assert float(nut.min()) >= 0
    z = torch.zeros(3, dtype=torch.float64)
    shear = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64)
    nut_s = O.eddy_viscosity(z, shear, z, z, ops)
    if model == "smagorinsky":
        expect = (0.17 * 0.1) ** 2 * np.sqrt(2 * 2 * (0.5 * shear.numpy()) ** 2 + 1e-30)
        assert np.allclose(nut_s.numpy(), expect, rtol=1e-12)
    else:
        assert float(nut_s.abs().max()) < 1e-12
    # float64 numpy restatement from the textbook formulas
    g11, g12, g21, g22 = (x.numpy() for x in g)
    S11, S22, S12 = g11, g22, 0.5 * (g12 + g21)
    ss = S11**2 + S22**2 + 2 * S12**2
    if model == "smagorinsky":
        ref = (0.17 * 0.1) ** 2 * np.sqrt(2 * ss + 1e-30)
    elif model == "wale":
        G = np.stack([[g11, g12], [g21, g22]])          # G[i][j] = du_i / dx_j
        GG = np.einsum("ikn,kjn->ijn", G, G)
        Sd = 0.5 * (GG + GG.transpose(1, 0, 2))
        tr = (GG[0, 0] + GG[1, 1]) / 3.0
        Sd[0, 0] -= tr
        Sd[1, 1] -= tr
        sd = Sd[0, 0] ** 2 + Sd[1, 1] ** 2 + 2 * Sd[0, 1] ** 2 + tr**2   # the zz entry is -tr
        ref = (0.5 * 0.1) ** 2 * sd**1.5 / (ss**2.5 + sd**1.25 + 1e-30)
    else:
        a = np.stack([[g11, g21], [g12, g22]])           # alpha_ij = du_j / dx_i
        beta = 0.01 * np.einsum("kin,kjn->ijn", a, a)
        bb = beta[0, 0] * beta[1, 1] - beta[0, 1] ** 2
        ref = 0.07 * np.sqrt(bb / (a**2).sum((0, 1)) + 1e-30)
        ref = ref  # c sqrt(B / alpha:alpha) with delta^2 inside beta
    assert np.allclose(nut.numpy(), ref, rtol=1e-9, atol=1e-15)


# ----------------------------------------------------------------- golden vectors
def test_masks_match_reference(golden):
    """Mask builder (signed distance, C1 step, window, force weights, thermal masks) equals
    the reference's tensors for the 256x128 disc and the 1024x512 heated airfoil."""
    F = golden("ns2d_masks.npz")
    for prefix, cfg, live in (("c1_", cases.c1_cfg("float64"), False), ("c2_", cases.c2_cfg(dtype="float64"), True)):
        ops = _ops(cfg)
        g = _geom(cfg, ops, live=live)
        assert tuple(g.window) == tuple(F[prefix + "window"])
        for k in ("chi_u", "chi_v", "chi_c", "a_u", "a_v", "k_u", "k_v"):
            assert np.abs(getattr(g, k).numpy() - F[prefix + k]).max() < 1e-14, (prefix, k)
        if prefix == "c2_":
            for k in ("chi_t", "a_t", "w_t", "vol_u", "vol_v", "ins_u", "ins_v"):
                assert np.abs(getattr(g.thermal, k).numpy() - F[prefix + "th_" + k]).max() < 1e-14, k
            w = ops.phys.groups[0].weight.numpy()
            j0, j1, i0, i1 = F["c2_dye_box"]
            assert np.abs(w[j0:j1, i0:i1] - F["c2_dye_weight"]).max() < 1e-14
            outside = w.copy()
            outside[j0:j1, i0:i1] = 0.0
            assert not outside.any()


def test_oracle_traced_step_matches_reference(golden):
    """One traced float64 step (every stage field) and 20 more steps equal the reference's."""
    F = golden("ns2d_trace_64x32.npz")
    cfg = cases.trace_cfg("float64")
    ops = _ops(cfg)
    geom = _geom(cfg, ops)
    tr = {}
    new, c = O.step_full(_state(F["u0"], F["v0"], ops), geom, ops, trace=tr)
    cu = lambda u: 0.5 * (u[:, 1:] + u[:, :-1])  # noqa: E731
    cv = lambda v: 0.5 * (v[1:] + v[:-1])  # noqa: E731
    for s in (1, 2, 3):
        for kind in ("explicit", "penalised", "projected"):
            assert np.abs(cu(tr[f"stage{s}/{kind}_u"].numpy()) - F[f"trace:stage{s}/{kind}_u"]).max() < 1e-12
            assert np.abs(cv(tr[f"stage{s}/{kind}_v"].numpy()) - F[f"trace:stage{s}/{kind}_v"]).max() < 1e-12
        assert np.abs(tr[f"stage{s}/divergence"].numpy() - F[f"trace:stage{s}/divergence"]).max() < 1e-10
        assert np.abs(tr[f"stage{s}/pressure_potential"].numpy() -
                      F[f"trace:stage{s}/pressure_potential"]).max() < 1e-12
    assert np.abs(new.p.numpy() - F["p1"]).max() < 1e-12
    assert np.abs(c.numpy() - F["coeffs1"]).max() < 1e-14
    cur = new
    for _ in range(20):
        cur, c = O.step_full(cur, geom, ops)
    assert np.abs(cur.u.numpy() - F["u21"]).max() < 1e-13
    assert np.abs(c.numpy() - F["coeffs_hist"][-1]).max() < 1e-13


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_oracle_heated_airfoil_matches_reference(golden, dtype):
    """128x64 heated airfoil, 20 steps: u, v, p, theta, dye and heat outputs."""
    tag = "f64" if dtype == "float64" else "f32"
    F = golden(f"ns2d_c2mini_{tag}.npz")
    cfg = cases.c2_cfg(128, 64, dtype)
    ops = _ops(cfg)
    geom = _geom(cfg, ops, live=True)
    cur = _state(F["u0"], F["v0"], ops, theta=F["theta0"], dye=F["dye0"])
    f64 = dtype == "float64"
    for n in range(1, 21):
        out = {}
        cur, c = O.step_full(cur, geom, ops, out=out)
        assert np.abs(c.numpy()[0] - F["hist_coeffs"][n - 1]).max() < (1e-12 if f64 else 1e-5)
        assert np.abs(out["heat"].numpy() - F["hist_heat"][n - 1]).max() < (1e-13 if f64 else 1e-6)
    for k in ("u", "v", "theta", "dye"):
        assert np.abs(getattr(cur, k).numpy() - F[f"{k}20"]).max() < (1e-13 if f64 else 2e-6), k
    assert np.abs(cur.p.numpy() - F["p20"]).max() < (1e-12 if f64 else 2e-5)


def test_oracle_c1_early_steps_match_reference(golden):
    """256x128 disc case, float32, first 10 steps (the full 500 run is a GPU-test bar)."""
    F = golden("ns2d_c1_f32.npz")
    cfg = cases.c1_cfg("float32")
    ops = _ops(cfg)
    geom = _geom(cfg, ops)
    cur = _state(F["u0"], F["v0"], ops)
    for _ in range(10):
        cur, c = O.step_full(cur, geom, ops)
    assert np.abs(cur.u.numpy() - F["u10"]).max() < 2e-6
    assert np.abs(cur.p.numpy() - F["p10"]).max() < 2e-5
    assert np.abs(np.asarray(F["dt_hist"][:10]) - F["dt_hist"][:10]).max() == 0


def test_oracle_wrong_stage_order_is_caught(golden):
    """A planted mistake (buoyancy before the penalisation) moves the result off the
    reference's by far more than round-off: the fixture pins the stage order."""
    F = golden("ns2d_c2mini_f64.npz")
    cfg = cases.c2_cfg(128, 64, "float64")
    ops = _ops(cfg)
    geom = _geom(cfg, ops, live=True)
    orig = O.add_buoyancy

    def swapped(uh, vh, theta, geom_, ops_, ph, bdt, w, acc_u, acc_v):
        # re-penalise after adding buoyancy: the solid keeps none of it
        uh, vh, acc_u, acc_v = orig(uh, vh, theta, geom_, ops_, ph, bdt, w, acc_u, acc_v)
        uh, vh, _, _ = O.penalise(uh, vh, geom_, ops_, bdt)
        return uh, vh, acc_u, acc_v

    O.add_buoyancy = swapped
    try:
        cur = _state(F["u0"], F["v0"], ops, theta=F["theta0"], dye=F["dye0"])
        cur, _ = O.step_full(cur, geom, ops)
    finally:
        O.add_buoyancy = orig
    assert np.abs(cur.v.numpy() - F["v1"]).max() > 1e-8


def test_oracle_heat_flux_bodies_match_reference(golden):
    """Heat-flux bodies (a disc releasing G = 2 through its skin beside a disc held at theta = 1),
    float64, 20 steps: this package's thermal window arrays equal the reference's, and the
    oracle's u, v, p, theta and heat outputs (the held disc's heat, the flux disc's mean skin
    temperature, walls, storage, advective flow) follow the reference's to round-off."""
    F = golden("ns2d_flux_bodies.npz")
    cfg = cases.flux_cfg("float64")
    ops = _ops(cfg)
    geom = _geom(cfg, ops, live=True)
    assert tuple(geom.window) == tuple(int(x) for x in F["window"])
    tg = geom.thermal
    assert tuple(tg.flux_ids) == tuple(F["flux_ids"]) and tuple(tg.temp_ids) == tuple(F["temp_ids"])
    for k in ("q_skin", "w_f", "perimeter", "ins_u", "ins_v", "chi_t", "flux_g"):
        assert np.abs(getattr(tg, k).numpy() - F["g_" + k]).max() <= 1e-12 * max(1.0, np.abs(F["g_" + k]).max()), k
    cur = _state(F["u0"], F["v0"], ops, theta=F["theta0"])
    for n in range(1, 21):
        out = {}
        cur, c = O.step_full(cur, geom, ops, out=out)
        assert np.abs(c.numpy() - F["hist_coeffs"][n - 1]).max() < 1e-11
        for k in ("heat", "theta_wall", "wall_heat", "heat_storage", "flow_heat"):
            want = F["hist_" + k][n - 1]
            assert np.abs(out[k].numpy() - want).max() <= 1e-11 * max(1.0, np.abs(want).max()), (n, k)
    for k in ("u", "v", "theta"):
        assert np.abs(getattr(cur, k).numpy() - F[f"{k}20"]).max() < 1e-12, k
    assert np.abs(cur.p.numpy() - F["p20"]).max() < 1e-10


def test_heat_flux_skin_releases_the_prescribed_rate(golden):
    """What the mathematics fixes for the skin source: kappa * q_skin summed over the window
    cells times h^2 is the body's heat rate G P / Pe (the skin weights sum to one), the skin
    lies in the diffuse band of the flux body only, and a planted mistake (the source without
    kappa) breaks the reference comparison."""
    F = golden("ns2d_flux_bodies.npz")
    cfg = cases.flux_cfg("float64")
    ops = _ops(cfg)
    geom = _geom(cfg, ops, live=True)
    tg = geom.thermal
    kappa = float(ops.phys.kappa)
    rate = kappa * float(tg.q_skin.sum()) * ops.h ** 2
    b = tg.flux_ids[0]
    want = cfg.bodies[b].thermal_value * float(tg.perimeter[b]) / (cfg.re * cfg.thermal.prandtl)
    assert rate == pytest.approx(want, rel=1e-12)
    assert float(tg.w_f.sum()) == pytest.approx(1.0, rel=1e-14)
    assert float(tg.w_f.min()) >= 0.0
    # planted mistake: the skin source without kappa moves theta off the reference's at step 1
    good = O.step_full(_state(F["u0"], F["v0"], ops, theta=F["theta0"]), geom, ops)[0]
    assert np.abs(good.theta.numpy() - F["theta1"]).max() < 1e-12
    tg.q_skin = tg.q_skin / kappa
    bad = O.step_full(_state(F["u0"], F["v0"], ops, theta=F["theta0"]), geom, ops)[0]
    assert np.abs(bad.theta.numpy() - F["theta1"]).max() > 1e-3
