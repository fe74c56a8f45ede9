"""Train and judge the learned closure end to end on one GPU.

1. Fine DNS (default 2048^2; nu = nuv = 5e-5, hnu = hnuv = 0.5, Omega = 5,
   random forcing on kf = 40 for psi and vz at injection 1.0, dt = 2e-4 fixed)
   from a seeded random-phase start: spin-up, then a training window with
   `snapshots` coarse-grained snapshots (inputs and subgrid targets, see
   rot2d3c_amd.closure), then an evaluation window whose coarse-grained
   spectra and fluxes are averaged in time.
2. Training of the closure network on the GPU (90% / 10% split).
3. Coarse runs (default 256^2, same parameters, fixed dt = dt nf/nc: the
   DNS's CFL number) over the evaluation window, started from the
   coarse-grained DNS state at its start:
   the plain coarse solver and the closure run.  Both are judged by their
   time-averaged energy spectrum and flux against the coarse-grained DNS for
   k <= n_coarse/3, and timed in GPU-seconds per unit simulated time against
   the DNS.

Usage: python -m rot2d3c_amd.closure_study --out DIR [--nf 2048 --nc 256 ...]
Writes DIR/closure_params.npz (parameters and normalisation) and
DIR/closure_study.json, printing a progress line per phase.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from . import closure as L
from . import ic
from . import spectral as S


class Forcing:
    """rand_force draws of one run, stepped: the pre-loop pair advances both
    seeds, then each step draws psi (seed) and vz (seed+1) forcing
    (src/Rot2D3C.para.f90:377-380,413-415)."""

    def __init__(self, kf, f0, seed):
        self.kf, self.f0 = kf, f0
        self.s0, self.s1 = seed, seed + 1
        self._draw()

    def _draw(self):
        kx, ky, ph, self.s0 = S.rand_force_draw(self.kf, self.s0)
        a = S.ForceDraw(kx, ky, ph, self.f0, self.kf)
        kx, ky, ph, self.s1 = S.rand_force_draw(self.kf, self.s1)
        b = S.ForceDraw(kx, ky, ph, self.f0, 1.0)
        return a, b

    def take(self, k):
        return [self._draw() for _ in range(k)]


def log(*a):
    print(*a, flush=True)


def spectra(ps, vz):
    ek, ezk = S.spectrum(ps, vz)
    _, (x0, x1, x2) = S.transfers(ps, vz)
    return np.asarray(ek), np.asarray(ezk), np.asarray(x1), np.asarray(x2)


def errors(les, dns, kmax):
    """Relative spectrum errors over k = 1..kmax and the flux error
    (max |difference| over the same range / max |DNS flux|)."""
    ek, ezk, fx1, fx2 = les
    dk, dzk, dx1, dx2 = dns
    r = slice(0, kmax)
    e_tot = (ek + ezk)[r]
    d_tot = (dk + dzk)[r]
    rel = np.abs(e_tot - d_tot) / d_tot
    flux = np.max(np.abs(fx1[r] - dx1[r])) / max(np.max(np.abs(dx1[r])), 1e-300)
    return {"spectrum_rel_err_max": float(rel.max()), "spectrum_rel_err_mean": float(rel.mean()),
            "spectrum_rel_err_k_le_40_max": float(rel[:40].max()),
            "flux_err": float(flux), "per_k": [float(x) for x in rel]}


def timed_steps(ps0, vz0, p, dt, forcing, steps, net=None, sin=None, sout=None):
    """Seconds per step of a one-run ensemble (each step one replay of the
    captured step graph, the closure inside it when net is given), timed over
    one run() call of `steps` steps after a warm-up call; host-side table
    building is outside the timed region."""
    from . import ensemble as E
    ps, vz = ps0.clone(), vz0.clone()
    ens = E.Ensemble([E.Member(ps, vz, p, dt)], schedules="auto")
    if net is not None:
        ens.set_closure(0, net, sin, sout)
    warm = ens.forcing_array([[pair] for pair in forcing.take(20)])
    table = ens.forcing_array([[pair] for pair in forcing.take(steps)])
    ens.run(20, warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps, table)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    ens.close()
    return sec


def coarse_run(ps0, vz0, p, t_end, dt, forcing, net=None, sin=None, sout=None,
               samples=200, energy_bound=None, chunk=50):
    """Advance a copy of (ps0, vz0) to simulated time t_end with fixed dt as a
    one-run ensemble, the closure inside each step's graph when net is given.
    Every `chunk` steps the energy is checked; a closure run that stops being
    finite or exceeds energy_bound restarts the chunk from its last good state
    as the plain solver (the fallback).  Returns (time-averaged statistics,
    steps, finite, fallbacks)."""
    from . import ensemble as E
    ps, vz = ps0.clone(), vz0.clone()
    ens = E.Ensemble([E.Member(ps, vz, p, dt)], schedules="auto")
    if net is not None:
        ens.set_closure(0, net, sin, sout)
    nsteps = int(round(t_end / dt))
    every = max(1, nsteps // samples)
    acc, nsamp, steps, finite, fallbacks = None, 0, 0, True, 0
    good = (ps.clone(), vz.clone())
    closure_on = net is not None
    while steps < nsteps:
        c = min(chunk, nsteps - steps)
        ens.run(c, [[pair] for pair in forcing.take(c)])
        e = S.energy(ps, 1) + S.energy(vz, 0)
        bad = not math.isfinite(e) or (energy_bound is not None and e > energy_bound)
        if bad and closure_on:
            ps.copy_(good[0])
            vz.copy_(good[1])
            ens.set_closure(0, None)
            closure_on = False
            fallbacks += 1
            continue
        if not math.isfinite(e):
            finite = False
            break
        good[0].copy_(ps)
        good[1].copy_(vz)
        steps += c
        if steps // every > (steps - c) // every:
            s = spectra(ps, vz)
            acc = [np.array(x) for x in s] if acc is None else [u + v for u, v in zip(acc, s)]
            nsamp += 1
    ens.close()
    stats = None if acc is None else [u / nsamp for u in acc]
    return stats, steps, finite, fallbacks


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--nf", type=int, default=2048)
    ap.add_argument("--nc", type=int, default=256)
    ap.add_argument("--dt", type=float, default=2e-4)
    ap.add_argument("--spinup", type=int, default=15000)
    ap.add_argument("--train-steps", type=int, default=25000)
    ap.add_argument("--snapshots", type=int, default=1000)
    ap.add_argument("--eval-steps", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--dt-coarse", type=float, default=None,
                    help="fixed dt of the coarse runs (default dt * nf/nc: the DNS's CFL number)")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nf, nc = a.nf, a.nc
    kf, f0 = 40.0, 1.0
    p = S.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    ps0, vz0 = ic.random_phase_ic(nf, seed=1234 + a.seed)
    ps = torch.from_numpy(ps0).to(dev)
    vz = torch.from_numpy(vz0).to(dev)
    forcing = Forcing(kf, f0, a.seed)
    out = {"config": vars(a), "dns": {}, "train": {}, "coarse": {}}

    def advance(k):
        done = 0
        while done < k:
            c = min(500, k - done)
            S.rk4_run(ps, vz, p, forcing.take(c), c, dt=a.dt)
            done += c

    # 1. spin-up
    t0 = time.perf_counter()
    energies = []
    for i in range(0, a.spinup, 2500):
        advance(min(2500, a.spinup - i))
        e = (S.energy(ps, 1), S.energy(vz, 0))
        energies.append(e)
        log("dns spin-up step %d energy %.4f %.4f" % (i + min(2500, a.spinup - i), *e))
    out["dns"]["spinup_energy"] = energies
    # DNS speed: a timed chunk without diagnostics
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    advance(1000)
    torch.cuda.synchronize()
    dns_ms = (time.perf_counter() - t1) * 1e3 / 1000
    out["dns"]["ms_per_step"] = dns_ms
    out["dns"]["gpu_s_per_time"] = dns_ms * 1e-3 / a.dt
    log("dns %.4f ms/step, %.3f GPU-s per unit time" % (dns_ms, out["dns"]["gpu_s_per_time"]))

    # 2. training window: snapshots
    every = max(1, a.train_steps // a.snapshots)
    X = torch.empty((a.snapshots, nc, nc, 3), dtype=torch.float32, device=dev)
    Y = torch.empty((a.snapshots, nc, nc, 2), dtype=torch.float32, device=dev)
    for k in range(a.snapshots):
        advance(every)
        _, _, x, y = L.snapshot(ps, vz, nc)
        X[k].copy_(x)
        Y[k].copy_(y)
        if k % 100 == 0:
            log("snapshot %d/%d" % (k, a.snapshots))
    torch.cuda.synchronize()
    sx, sy = L.channel_std(X), L.channel_std(Y)
    for c in range(3):
        X[..., c] *= 1.0 / sx[c]
    for c in range(2):
        Y[..., c] *= 1.0 / sy[c]
    nval = max(a.batch, a.snapshots // 10)
    Xtr, Ytr = X[:-nval].contiguous(), Y[:-nval].contiguous()
    Xva, Yva = X[-nval:].contiguous(), Y[-nval:].contiguous()

    # 3. training
    net = L.Net(a.batch, nc, nc, dev, seed=a.seed)
    t2 = time.perf_counter()
    hist = L.train(net, Xtr, Ytr, a.epochs, lr=a.lr, lr_end=a.lr / 10, seed=a.seed, log=log)
    torch.cuda.synchronize()
    out["train"]["seconds"] = time.perf_counter() - t2
    out["train"]["loss"] = hist
    out["train"]["val_loss"] = L.evaluate(net, Xva, Yva)
    out["train"]["val_loss_zero_model"] = float((Yva.double() ** 2).mean().item())
    log("train loss %.4f val %.4f (zero model %.4f), %.1f s" % (
        hist[-1], out["train"]["val_loss"], out["train"]["val_loss_zero_model"],
        out["train"]["seconds"]))
    params = net.params().cpu().numpy()
    net.close()
    del X, Y, Xtr, Ytr
    torch.cuda.empty_cache()
    scale_in = [1.0 / v for v in sx]
    np.savez(os.path.join(a.out, "closure_params.npz"), params=params, scale_in=np.array(scale_in),
             scale_out=np.array(sy), nf=nf, nc=nc)

    # 4. evaluation window: coarse-grained DNS statistics
    ps_c0 = L.coarse_grain(ps, nc)
    vz_c0 = L.coarse_grain(vz, nc)
    acc, nsamp = None, 0
    samp = max(1, a.eval_steps // 200)
    for k in range(0, a.eval_steps, samp):
        advance(samp)
        s = spectra(L.coarse_grain(ps, nc), L.coarse_grain(vz, nc))
        acc = [np.array(x) for x in s] if acc is None else [u + v for u, v in zip(acc, s)]
        nsamp += 1
    dns_stats = [u / nsamp for u in acc]
    T = a.eval_steps * a.dt
    log("dns evaluation window done (%d samples over t = %.3f)" % (nsamp, T))
    del ps, vz
    torch.cuda.empty_cache()

    # 5. coarse runs over the same window
    kmax = nc // 3
    infer = L.Net(1, nc, nc, dev, params=params)
    e0 = S.energy(ps_c0, 1) + S.energy(vz_c0, 0)
    dtc = a.dt * nf / nc if a.dt_coarse is None else a.dt_coarse
    for name, net in (("plain", None), ("closure", infer)):
        stats, steps, finite, fb = coarse_run(
            ps_c0, vz_c0, p, T, dtc, Forcing(kf, f0, a.seed + 7), net, scale_in, sy,
            energy_bound=20.0 * e0)
        sec = timed_steps(ps_c0, vz_c0, p, dtc, Forcing(kf, f0, a.seed + 11), 500, net, scale_in,
                          sy)
        r = {"finite": finite, "steps": steps, "dt": dtc, "sim_time": steps * dtc,
             "ms_per_step": sec * 1e3, "gpu_s_per_time": sec / dtc, "fallbacks": fb}
        if finite and stats is not None:
            r.update(errors(stats, dns_stats, kmax))
            r["spectrum"] = [float(x) for x in (stats[0] + stats[1])[:kmax]]
        out["coarse"][name] = r
        log("%s: finite %s, %d steps, t %.3f, %.3f GPU-s/time, spectrum err max %s mean %s" % (
            name, finite, steps, r["sim_time"], r["gpu_s_per_time"], r.get("spectrum_rel_err_max"),
            r.get("spectrum_rel_err_mean")))
    out["dns"]["spectrum"] = [float(x) for x in (dns_stats[0] + dns_stats[1])[:kmax]]
    c = out["coarse"]["closure"]
    out["speedup_gpu_s_per_time"] = out["dns"]["gpu_s_per_time"] / c["gpu_s_per_time"]
    out["wall_seconds"] = time.perf_counter() - t0
    with open(os.path.join(a.out, "closure_study.json"), "w") as f:
        json.dump(out, f, indent=1)
    log("speedup %.1fx; wrote %s" % (out["speedup_gpu_s_per_time"], a.out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
