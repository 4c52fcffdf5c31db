"""Train the closure network on filtered fine-grid solves and evaluate it by the spectrum of
coarse solves (``python -m aerolume_amd.closure.train``; prints one JSON report).

1. Fine solves (no model) of decaying box turbulence from seeded spectral fields; snapshots are
   face-filtered to the coarse grid with their subgrid stresses.
2. Inputs: coarse cell-centred velocity gradients; target: the eddy viscosity matching the
   subgrid energy transfer (``data.sgs_target``); the network is trained by its own HIP
   forward / backward / Adam kernels (MSE on the scaled target).
3. Coarse solves from the filtered state of a second seed with the learned closure (WALE
   fallback), with WALE alone and with no model; their kinetic-energy spectra at the end time
   are compared with the filtered fine spectrum at the same time; wall times are compared
   with the fine solve over the same interval.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from ..ns3d import solver as S3
from . import data as D
from .model import LearnedClosure
from .network import CNN3D


def run_fine(n: int, r: int, seed: int, t_start: float, t_end: float, n_snap: int, re: float):
    """Fine solve; returns the filtered snapshots [(t, uc, vc, wc, tau)] at n_snap times in
    [t_start, t_end] and the fine wall time from t_start to t_end."""
    fs = S3.FlowSolver3D(D.box_config(n, re=re))
    u, v, w = D.to_staggered(D.spectral_velocity(n, seed), fs.ops)
    fs.set_velocity(u, v, w)
    times = np.linspace(t_start, t_end, n_snap)
    snaps, wall = [], 0.0
    for t_target in times:
        t0 = time.perf_counter()
        while fs.t < t_target - 1e-12:
            fs.advance(10)
        torch.cuda.synchronize()
        if t_target > times[0]:
            wall += time.perf_counter() - t0
        st = fs.state
        uc, vc, wc = D.face_filter(st.u, st.v, st.w, r)
        snaps.append((fs.t, uc, vc, wc, D.sgs_stress(st.u, st.v, st.w, r)))
    steps = fs.steps
    del fs
    torch.cuda.empty_cache()
    return snaps, wall, steps


def features_and_targets(snaps, m: int, re: float):
    ops = S3.make_operators3d(D.box_config(m, re=re))
    plan = S3.get_plan(ops, S3._empty_geometry(ops))
    xs, ys = [], []
    for _, uc, vc, wc, tau in snaps:
        g = torch.empty(9, m, m, m, device=ops.device, dtype=torch.float32)
        plan.cell_gradient(uc, vc, wc, g, 1.0)
        xs.append(g)
        ys.append(D.sgs_target(tau.float(), g))
    return xs, ys


def train(xs, ys, m: int, iters: int, lr: float, seed: int):
    g_scale = float(torch.sqrt(torch.stack([(x * x).mean() for x in xs]).mean() * 9.0))
    nu_all = torch.cat([y.reshape(-1) for y in ys])
    nu_scale = float(nu_all.mean()) + 1e-12
    trust = float(torch.quantile(nu_all[::max(1, nu_all.numel() // 1_000_000)], 0.999)) * 4.0
    net = CNN3D((m, m, m), seed=seed)
    xn = [x / g_scale for x in xs]
    yn = [y / nu_scale for y in ys]

    def data_loss():
        return float(torch.stack([((net.infer(x)[0] - y) ** 2).mean() for x, y in zip(xn, yn)]).mean())

    losses = [data_loss()]
    for it in range(iters):
        j = it % len(xn)
        out = net.forward(xn[j])
        diff = out[0] - yn[j]
        net.backward((2.0 / diff.numel()) * diff[None].contiguous())
        net.adam_step(lr=lr if it < iters // 2 else lr * 0.3)
    losses.append(data_loss())
    # a priori: correlation of prediction and target over all snapshots
    pred = torch.cat([net.infer(x)[0].reshape(-1) for x in xn])
    targ = torch.cat([y.reshape(-1) for y in yn])
    corr = float(torch.corrcoef(torch.stack([pred, targ]))[0, 1])
    # the same statistic for WALE, the model the network stands in for
    wale = torch.cat([D.wale_nut(x).reshape(-1) for x in xs])
    corr_wale = float(torch.corrcoef(torch.stack([wale, targ]))[0, 1])
    return net, g_scale, nu_scale, trust, losses, corr, corr_wale


def run_coarse(m: int, start, t_end: float, re: float, closure=None, turbulence: str = "wale",
               advection: str = "upwind3"):
    t0_sim, uc, vc, wc = start
    fs = S3.FlowSolver3D(D.box_config(m, re=re, turbulence=turbulence, advection=advection), closure=closure)
    fs.set_velocity(uc, vc, wc)
    fs._state.t.fill_(t0_sim)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while fs.t < t_end - 1e-12:
        fs.advance(10)
        if not bool(torch.isfinite(fs.state.u[0, 0, :4]).all()):
            break
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = fs.state
    ok = bool(torch.isfinite(st.u).all())
    k, e = D.kinetic_energy_spectrum(st.u, st.v, st.w)
    return {"wall": wall, "steps": fs.steps, "t": fs.t, "finite": ok, "k": k, "E": e}


def spectrum_error(e, e_ref, kmax: int) -> float:
    kk = slice(1, kmax + 1)
    return float(np.max(np.abs(e[kk] - e_ref[kk]) / e_ref[kk]))


def experiment(n: int = 256, r: int = 4, re: float = 1000.0, t_start: float = 0.5, t_end: float = 1.5,
               n_snap: int = 6, iters: int = 300, lr: float = 2e-3, train_seed: int = 505, test_seed: int = 606,
               coarse_advection: str = "upwind3"):
    m = n // r
    t0 = time.perf_counter()
    snaps, _, _ = run_fine(n, r, train_seed, t_start, t_end, n_snap, re)
    xs, ys = features_and_targets(snaps, m, re)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, g_scale, nu_scale, trust, losses, corr, corr_wale = train(xs, ys, m, iters, lr, seed=0)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del snaps, xs, ys
    test, fine_wall, fine_steps = run_fine(n, r, test_seed, t_start, t_end, 2, re)
    start = test[0][:4]
    _, ue, ve, we, _ = test[-1]
    k, e_ref = D.kinetic_energy_spectrum(ue, ve, we)
    closure = LearnedClosure(net, g_scale, nu_scale, trust)
    adv = coarse_advection
    runs = {"learned": run_coarse(m, start, t_end, re, closure=closure, advection=adv),
            "wale": run_coarse(m, start, t_end, re, turbulence="wale", advection=adv),
            "none": run_coarse(m, start, t_end, re, turbulence="none", advection=adv)}
    kc = m // 2
    report = {"fine_grid": n, "coarse_grid": m, "filter": r, "re": re, "t_start": t_start, "t_end": t_end,
              "coarse_advection": adv,
              "train_snapshots": n_snap, "train_iters": iters, "train_seconds": round(t_train, 2),
              "data_seconds": round(t_data, 2), "a_priori_correlation": round(corr, 4),
              "a_priori_correlation_wale": round(corr_wale, 4),
              "loss_first_last": [round(losses[0], 5), round(losses[-1], 5)],   # mean over all snapshots
              "g_scale": g_scale, "nu_scale": nu_scale, "trust": trust,
              "fine_wall_s": round(fine_wall, 3), "fine_steps": fine_steps, "k_c": kc,
              "E_filtered_fine": [float(x) for x in e_ref[:kc + 1]]}
    for name, rr in runs.items():
        if not rr["finite"]:
            report[name] = {"finite": False, "wall_s": round(rr["wall"], 3), "steps": rr["steps"]}
            continue
        report[name] = {"wall_s": round(rr["wall"], 3), "steps": rr["steps"], "finite": rr["finite"],
                        "max_rel_spectrum_error_k_le_kc": round(spectrum_error(rr["E"], e_ref, kc), 4),
                        "max_rel_spectrum_error_k_le_kc_half": round(spectrum_error(rr["E"], e_ref, kc // 2), 4),
                        "speedup_vs_fine": round(fine_wall / max(rr["wall"], 1e-9), 2),
                        "E": [float(x) for x in rr["E"][:kc + 1]]}
    return report


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=256)
    ap.add_argument("--ratio", type=int, default=4)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--snapshots", type=int, default=6)
    ap.add_argument("--t-end", type=float, default=1.5)
    ap.add_argument("--coarse-advection", default="upwind3", choices=("upwind3", "central"))
    args = ap.parse_args()
    rep = experiment(n=args.fine, r=args.ratio, iters=args.iters, n_snap=args.snapshots, t_end=args.t_end,
                     coarse_advection=args.coarse_advection)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
