"""Train and judge the learned closure end to end on one GPU.

1. Fine DNS (default 2048^2; nu = nuv = 5e-5, hnu = hnuv = 0.5, Omega = 5,
   random forcing on kf = 40 for psi and vz at injection 1.0, dt = 2e-4 fixed)
   from a seeded random-phase start: spin-up, then a training window with
   `snapshots` coarse-grained snapshots (inputs and subgrid targets, see
   rot2d3c_amd.closure), then a tuning window and an evaluation window whose
   coarse-grained spectra and fluxes are averaged in time.
2. Training of the closure network on the GPU (90% / 10% split).
3. Coarse runs (default 256^2, same parameters, fixed dt = dt nf/nc: the
   DNS's CFL number), each as several forcing realisations averaged,
   started from the coarse-grained DNS state at the start of a window.  The
   DNS of each tuning or evaluation window is as many realisations from the
   window's start, forced by the same draws as the coarse runs (the first
   continues into the next window); the judge compares the means and takes
   the realisations' spread as the noise floor per shell.  On
   the tuning window each candidate low-k cutoff of the closure forcing is
   scored; on the evaluation window the plain coarse solver and the closure
   run with the chosen cutoff are judged by their time-averaged energy
   spectrum and flux against the coarse-grained DNS for k <= n_coarse/3, and
   timed in GPU-seconds per unit simulated time against the DNS.

Usage: python -m rot2d3c_amd.closure_study --out DIR [--nf 2048 --nc 256 ...]
Writes DIR/closure_params.npz (parameters and normalisation),
DIR/closure_study.json (everything) and DIR/closure_eval.json (the compact
record), printing a progress line per phase.
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

    def __init__(self, kf, f0, seed, n_coarse=None):
        """n_coarse: coarse-grain every draw for a coarse run of that size
        (closure.filtered_draw), so the coarse run is forced by G f."""
        self.kf, self.f0, self.nc = kf, f0, n_coarse
        self.s0, self.s1 = seed, seed + 1
        self._draw()

    def _draw(self):
        kx, ky, ph, self.s0 = S.rand_force_draw(self.kf, self.s0)
        a = S.ForceDraw(kx, ky, ph, self.f0, self.kf)
        kx, ky, ph, self.s1 = S.rand_force_draw(self.kf, self.s1)
        b = S.ForceDraw(kx, ky, ph, self.f0, 1.0)
        if self.nc is not None:
            a, b = L.filtered_draw(a, self.nc), L.filtered_draw(b, self.nc)
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


def timed_steps(ps0, vz0, p, dt, forcing, steps, net=None, sin=None, sout=None, kmin=0.0,
                kclip=0.0, clamp=0.0, sign_clip=0.0, period=1):
    """Seconds per step of a one-run ensemble (each step one replay of the
    captured step graph, the closure inside it when net is given), timed over
    one run() call of `steps` steps after a warm-up call; host-side table
    building is outside the timed region."""
    from . import ensemble as E
    ps, vz = ps0.clone(), vz0.clone()
    ens = E.Ensemble([E.Member(ps, vz, p, dt)], schedules=None)  # the default kernels
    if net is not None:
        net.set_output_clamp(clamp)
        net.set_sign_clip(sign_clip)
        ens.set_closure(0, net, sin, sout, kmin, kclip)
        ens.set_closure_period(period)
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


def coarse_runs(ps0, vz0, p, t_end, dt, seeds, kf, f0, params=None, sin=None, sout=None,
                kmin=0.0, samples=200, energy_bound=None, chunk=50, kclip=0.0,
                enstrophy_bound=None, trace=None, clamp=0.0, cfl_bound=None, sign_clip=0.0,
                safe=None, period=1, gain=None):
    """len(seeds) coarse runs from (ps0, vz0), one coarse-grained forcing
    realisation each (the filtered equation's forcing G f),
    advanced together as one ensemble to simulated time t_end with fixed dt;
    with `params` every run carries its own closure network (forcing on
    |k| >= kmin) inside its step graph.  Every `chunk` steps each run's energy
    and enstrophy are checked: a closure run that stops being finite or
    exceeds energy_bound or enstrophy_bound, or whose fixed dt exceeds
    cfl_bound times the CFL condition's dt, restarts the chunk from its last
    good state as the plain solver (the fallback).  The networks' outputs
    are saturated at +-clamp (0: none) and sign-clipped with strength
    sign_clip (closure.Net.set_sign_clip), and the closure forcing is scaled
    per shell by `gain` (closure.Net.set_forcing_gain; None: none).  safe =
    (cfl_on, cfl_off): a
    closure run whose fixed dt's CFL number exceeds cfl_on continues with the
    fully sign-clipped (damping only) network until it falls below cfl_off
    (checked every chunk; coarse_runs.safe_fraction is the fraction of
    run-chunks spent so).  Returns (statistics averaged over time and runs,
    steps, finite, fallbacks)."""
    from . import ensemble as E
    R = len(seeds)
    runs = [(ps0.clone(), vz0.clone()) for _ in range(R)]
    # the default schedule: the same kernels on every GPU, so the study is
    # reproducible bit for bit (an autotuned choice may pick kernels that
    # differ in rounding, and the coarse runs are chaotic)
    ens = E.Ensemble([E.Member(a, b, p, dt) for a, b in runs], schedules=None)
    nets = []
    if params is not None:
        for r in range(R):
            nets.append(L.Net(1, S._spec_n(ps0), S._spec_n(ps0), ps0.device, params=params))
            nets[-1].set_output_clamp(clamp)
            nets[-1].set_sign_clip(sign_clip)
            nets[-1].set_forcing_gain(gain)
            ens.set_closure(r, nets[-1], sin, sout, kmin, kclip)
        ens.set_closure_period(period)
    forcings = [Forcing(kf, f0, sd, S._spec_n(ps0)) for sd in seeds]
    nsteps = int(round(t_end / dt))
    every = max(1, nsteps // samples)
    accs, nsamp, steps, finite, fallbacks, cfl_max = [None] * R, 0, 0, True, 0, 0.0
    good = [(a.clone(), b.clone()) for a, b in runs]
    on = [params is not None] * R
    sm = None if safe is None else L.SafeMode(safe[0], safe[1], sign_clip)

    while steps < nsteps:
        c = min(chunk, nsteps - steps)
        draws = [f.take(c) for f in forcings]
        ens.run(c, [[draws[r][k] for r in range(R)] for k in range(c)])
        redo = False
        cfls = [float("nan")] * R
        for r, (a, b) in enumerate(runs):
            e = S.energy(a, 1) + S.energy(b, 0)
            z = S.energy(a, 2) + S.energy(b, 1)
            cfl = L.cfl_number(a, b, p, dt) if math.isfinite(e) else float("nan")
            bad = (not math.isfinite(e) or not math.isfinite(z)
                   or (energy_bound is not None and e > energy_bound)
                   or (enstrophy_bound is not None and z > enstrophy_bound)
                   or (cfl_bound is not None and cfl > cfl_bound))
            cfls[r] = cfl
            if on[r] and math.isfinite(cfl):
                cfl_max = max(cfl_max, cfl)
            if trace is not None:
                trace("step %d run %d closure %d%s energy %.6g enstrophy %.6g cfl %.3f%s" % (
                    steps + c, r, on[r], " safe" if sm and sm.active.get(r) else "", e, z, cfl,
                    " BAD" if bad else ""))
            if bad and on[r]:
                on[r] = False
                ens.set_closure(r, None)
                fallbacks += 1
                redo = True
            elif not math.isfinite(e):
                finite = False
        if not finite:
            break
        if redo:  # restart the chunk from the last good states
            for r, (a, b) in enumerate(runs):
                a.copy_(good[r][0])
                b.copy_(good[r][1])
            continue
        for r, (a, b) in enumerate(runs):
            good[r][0].copy_(a)
            good[r][1].copy_(b)
            if on[r] and sm is not None and sm.check(r, cfls[r]):
                sm.apply(nets[r], sm.active[r])
                ens.set_closure(r, nets[r], sin, sout, kmin, kclip)  # recaptures the step graph
        steps += c
        if steps // every > (steps - c) // every:
            for r, (a, b) in enumerate(runs):
                s_ = spectra(a, b)
                accs[r] = ([np.array(x) for x in s_] if accs[r] is None
                           else [u + v for u, v in zip(accs[r], s_)])
            nsamp += 1
    ens.close()
    per_run = None if accs[0] is None else [[u / nsamp for u in acc] for acc in accs]
    stats = None if per_run is None else [sum(x) / R for x in zip(*per_run)]
    coarse_runs.per_run = per_run  # each run's time-averaged statistics
    coarse_runs.cfl_max = cfl_max  # largest CFL number of a closure run at a check
    coarse_runs.safe_fraction = 0.0 if sm is None else sm.fraction
    return stats, steps, finite, fallbacks


def dns_window(advance, ps, vz, nc, steps, samples=200):
    """Advance the DNS `steps` steps, averaging the coarse-grained spectra and
    fluxes over `samples` instants."""
    acc, nsamp = None, 0
    samp = max(1, steps // samples)
    for _ in range(0, steps, samp):
        advance(samp)
        s_ = spectra(L.coarse_grain(ps, nc), L.coarse_grain(vz, nc))
        acc = [np.array(x) for x in s_] if acc is None else [u + v for u, v in zip(acc, s_)]
        nsamp += 1
    return [u / nsamp for u in acc]


def dns_realisations(ps, vz, p, dt, nc, steps, seeds, kf, f0, samples=200):
    """From the state (ps, vz), one DNS run per forcing seed over `steps`
    steps; run r is forced by the draws of seeds[r], the same draws (not
    coarse-grained) as coarse run r of coarse_runs.  Returns each run's
    time-averaged coarse-grained statistics (dns_window) and leaves ps, vz at
    the end state of the first run."""
    start = (ps.clone(), vz.clone())
    out, end0 = [], None
    for r, sd in enumerate(seeds):
        ps.copy_(start[0])
        vz.copy_(start[1])
        f = Forcing(kf, f0, sd)

        def advance(k):
            done = 0
            while done < k:
                c = min(500, k - done)
                S.rk4_run(ps, vz, p, f.take(c), c, dt=dt)
                done += c

        out.append(dns_window(advance, ps, vz, nc, steps, samples))
        if r == 0:
            end0 = (ps.clone(), vz.clone())
    ps.copy_(end0[0])
    vz.copy_(end0[1])
    return out


def mean_stats(runs):
    return [sum(x) / len(runs) for x in zip(*runs)]


def judge(les_runs, dns_runs, kmax, tol=0.05):
    """The closure's judge: the time-averaged energy spectrum E(k) = Ek + Ezk
    of the coarse runs (mean over their forcing realisations) against the
    mean of the coarse-grained DNS realisations forced by the same draws,
    for k = 1..kmax, and the 2D energy flux.  A shell agrees when its
    relative error is at most `tol`, or at most the noise floor 2 sigma of
    the difference of the two means (sigma from the realisations' spread),
    where that is larger.  The flux error is max_k |difference of the mean
    fluxes| / max_k |DNS flux|, and its noise floor 2 sigma of the largest
    difference of the means in the same units."""
    e_les = np.array([(s[0] + s[1])[:kmax] for s in les_runs])
    e_dns = np.array([(s[0] + s[1])[:kmax] for s in dns_runs])
    m_les, m_dns = e_les.mean(0), e_dns.mean(0)

    def se(e):
        return e.std(0, ddof=1) / math.sqrt(len(e)) if len(e) > 1 else np.zeros(e.shape[1])

    noise = 2.0 * np.sqrt(se(e_les) ** 2 + se(e_dns) ** 2) / m_dns
    rel = np.abs(m_les - m_dns) / m_dns
    f_les = np.array([s[2][:kmax] for s in les_runs])
    f_dns = np.array([s[2][:kmax] for s in dns_runs])
    fx_les, fx_dns = f_les.mean(0), f_dns.mean(0)
    fscale = max(np.max(np.abs(fx_dns)), 1e-300)
    flux = np.max(np.abs(fx_les - fx_dns)) / fscale
    flux_noise = 2.0 * np.max(np.sqrt(se(f_les) ** 2 + se(f_dns) ** 2)) / fscale
    ok = rel <= np.maximum(tol, noise)
    return {"spectrum_rel_err_mean": float(rel.mean()), "spectrum_rel_err_max": float(rel.max()),
            "flux_err": float(flux), "flux_noise_floor": float(flux_noise), "shells": int(kmax),
            "shells_within_5pct": int(np.sum(rel <= tol)),
            "shells_within_5pct_or_noise": int(np.sum(ok)),
            "shells_outside": [int(k + 1) for k in np.nonzero(~ok)[0]],
            "dns_realisations": len(dns_runs), "les_realisations": len(les_runs),
            "dns_spread_rel": [float(x) for x in se(e_dns) / m_dns],
            "noise_floor_rel": [float(x) for x in noise], "per_k": [float(x) for x in rel],
            "ratio": [float(x) for x in m_les / m_dns]}


def closure_settings(z):
    """The inference settings stored with a trained network (closure_params.npz)."""
    g = lambda k, d: float(z[k]) if k in z else d
    return dict(kmin=g("kmin", 0.0), kclip=g("kclip", 0.0), clamp=g("out_clamp", 0.0),
                sign_clip=g("sign_clip", 0.0), period=int(g("period", 1)),
                ktaper=g("ktaper", 0.0),
                safe=tuple(float(v) for v in z["safe_cfl"]) if "safe_cfl" in z else None)


def live_evaluation(npz_path, dev, spinup=15000, window=10000, realisations=2, samples=100,
                    dt=2e-4, seed=0, cfl_guard=3.0, guard=4.0, log=None):
    """Judge a trained closure on a DNS window generated here: the 2048^2 DNS
    of the study (same parameters and seeded start) spun up `spinup` steps,
    then `realisations` DNS runs over `window` steps from that state, forced
    by the same draws as the coarse runs; the plain 256^2 solver and the
    closure run (its stored inference settings, stabiliser and fallback
    guards as in the study), each as `realisations` runs from the
    coarse-grained start state.  Returns {"plain": judge(...), "closure":
    judge(...), ...}."""
    log = log or (lambda *a: None)
    z = np.load(npz_path)
    nf, nc = int(z["nf"]), int(z["nc"])
    cs = closure_settings(z)
    kf, f0 = 40.0, 1.0
    p = S.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    ps0, vz0 = ic.random_phase_ic(nf, seed=1234 + seed)
    ps = torch.from_numpy(ps0).to(dev)
    vz = torch.from_numpy(vz0).to(dev)
    del ps0, vz0
    forcing = Forcing(kf, f0, seed)
    done = 0
    while done < spinup:
        c = min(500, spinup - done)
        S.rk4_run(ps, vz, p, forcing.take(c), c, dt=dt)
        done += c
    dns_cfl = L.cfl_number(ps, vz, p, dt)
    log("closure judge: DNS %d^2 spun up %d steps (CFL number %.3f)" % (nf, spinup, dns_cfl))
    ps_c0, vz_c0 = L.coarse_grain(ps, nc), L.coarse_grain(vz, nc)
    e0 = S.energy(ps_c0, 1) + S.energy(vz_c0, 0)
    z0 = S.energy(ps_c0, 2) + S.energy(vz_c0, 1)
    seeds = [seed + 100 + r for r in range(realisations)]
    dns = dns_realisations(ps, vz, p, dt, nc, window, seeds, kf, f0, samples)
    log("closure judge: %d DNS realisations of %d steps" % (realisations, window))
    del ps, vz
    torch.cuda.empty_cache()
    dtc = dt * nf / nc
    kmax = nc // 3
    out = {"window_time": window * dt, "spinup_steps": spinup, "dns_cfl": dns_cfl,
           "coarse_dt": dtc, "settings": {k: v for k, v in cs.items() if k != "safe"}}
    safe = cs["safe"]
    for name, prm in (("plain", None), ("closure", z["params"])):
        stats, steps, finite, fb = coarse_runs(
            ps_c0, vz_c0, p, window * dt, dtc, seeds, kf, f0, prm, z["scale_in"], z["scale_out"],
            cs["kmin"], samples=samples, energy_bound=20.0 * e0, kclip=cs["kclip"],
            enstrophy_bound=guard * z0, clamp=cs["clamp"], cfl_bound=cfl_guard * dns_cfl,
            sign_clip=cs["sign_clip"], safe=safe, period=cs["period"],
            gain=L.taper_gain(cs["kmin"], cs["ktaper"], nc // 2 + 2) if cs["ktaper"] > 0 else None)
        r = {"finite": bool(finite), "fallbacks": fb, "steps": steps}
        if finite and stats is not None:
            r.update(judge(coarse_runs.per_run, dns, kmax))
        if prm is not None:
            r["safe_fraction"] = coarse_runs.safe_fraction
        out[name] = r
        log("closure judge: %s finite %s, spectrum err mean %s max %s, flux err %s, "
            "shells within 5%% %s, within 5%% or noise %s" % (
                name, finite, r.get("spectrum_rel_err_mean"), r.get("spectrum_rel_err_max"),
                r.get("flux_err"), r.get("shells_within_5pct"),
                r.get("shells_within_5pct_or_noise")))
    return out


def train_window(a, advance, ps, vz, out, log):
    """Training window of the DNS: coarse-grained snapshots, then training on
    the GPU.  Returns (params, scale_in, scale_out)."""
    dev, nc = ps.device, a.nc
    every = max(1, a.train_steps // a.snapshots)
    X = torch.empty((a.snapshots, nc, nc, 3), dtype=torch.float32, device=dev)
    Y = torch.empty((a.snapshots, nc, nc, 2), dtype=torch.float32, device=dev)
    for k in range(a.snapshots):
        advance(every)
        _, _, x, y = L.snapshot(ps, vz, nc, kmin=a.target_kmin)
        X[k].copy_(x)
        Y[k].copy_(y)
        if k % 100 == 0:
            log("snapshot %d/%d" % (k, a.snapshots))
    torch.cuda.synchronize()
    sx, sy = L.channel_std(X), L.channel_std(Y)
    if a.no_psi_input:  # the network sees laplacian(psi) and vz only
        sx[2] = math.inf
    for c in range(3):
        X[..., c] *= 1.0 / sx[c]
    for c in range(2):
        Y[..., c] *= 1.0 / sy[c]
    nval = max(a.batch, a.snapshots // 10)
    Xtr, Ytr = X[:-nval].contiguous(), Y[:-nval].contiguous()
    Xva, Yva = X[-nval:].contiguous(), Y[-nval:].contiguous()

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
    return params, scale_in, sy


def compact_record(out, command):
    """The evaluation record kept beside the trained network: DNS cost,
    training losses, the tuning scores, both coarse runs' errors and costs,
    the speedup and the per-shell spectrum ratios to the coarse-grained DNS."""
    keep = ("spectrum_rel_err_max", "spectrum_rel_err_mean", "spectrum_rel_err_k_le_40_max",
            "flux_err", "gpu_s_per_time", "ms_per_step", "dt", "realisations", "fallbacks", "cfl_max",
            "safe_fraction")
    dns = np.array(out["dns"]["spectrum"])
    rec = {"command": command,
           "dns": {"n": out["config"]["nf"], "ms_per_step": out["dns"]["ms_per_step"],
                   "gpu_s_per_time": out["dns"]["gpu_s_per_time"]},
           "train": {"val_loss": out["train"]["val_loss"],
                     "val_loss_zero_model": out["train"]["val_loss_zero_model"],
                     "seconds": out["train"]["seconds"],
                     "epochs": None if out["config"]["params"] else out["config"]["epochs"]},
           "kmin": out["kmin"], "ktaper": out["ktaper"], "gain": out["gain"],
           "kclip": out["kclip"], "clamp": out["clamp"],
           "sign_clip": out["sign_clip"], "period": out["period"],
           "selection": out.get("selection", "min tuning error"),
           "dns_cfl": out["dns"]["cfl"], "cfl_guard": out["config"]["cfl_guard"],
           "safe_cfl": out["config"]["safe_cfl"],
           "tune": {k: {m: v[m] for m in ("spectrum_rel_err_max", "spectrum_rel_err_mean",
                                           "flux_err", "shells_within_5pct",
                                           "shells_within_5pct_or_noise") if m in v}
                    for k, v in out["tune"].items()},
           "evaluation_window": out["config"]["eval_steps"] * out["config"]["dt"],
           "coarse_forcing": "coarse-grained draws (injection f0 G(k)^2)"}
    for name in ("plain", "closure"):
        rec[name] = {m: out["coarse"][name][m] for m in keep if m in out["coarse"][name]}
    for name in ("plain", "closure"):
        j = out["coarse"][name].get("judge")
        if j is not None:
            rec[name]["judge"] = {k: j[k] for k in (
                "spectrum_rel_err_mean", "spectrum_rel_err_max", "flux_err", "shells",
                "shells_within_5pct", "shells_within_5pct_or_noise", "shells_outside",
                "dns_realisations", "noise_floor_rel", "ratio")}
    rec["speedup_gpu_s_per_time"] = out["speedup_gpu_s_per_time"]
    rec["shells"] = int(dns.size)
    for name in ("closure", "plain"):
        sp = out["coarse"][name].get("spectrum")
        ratio = None if sp is None else np.array(sp) / dns
        rec["ratio_%s_to_dns" % name] = None if ratio is None else [float(x) for x in ratio]
        if name == "closure":
            rec["shells_within_5pct"] = 0 if ratio is None else int(np.sum(np.abs(ratio - 1) <= 0.05))
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--nf", type=int, default=2048)
    ap.add_argument("--nc", type=int, default=256)
    ap.add_argument("--dt", type=float, default=2e-4)
    ap.add_argument("--spinup", type=int, default=15000)
    ap.add_argument("--train-steps", type=int, default=25000)
    ap.add_argument("--snapshots", type=int, default=1000)
    ap.add_argument("--tune-steps", type=int, default=25000,
                    help="DNS steps of the window that picks the closure's low-k cutoff")
    ap.add_argument("--eval-steps", type=int, default=50000)
    ap.add_argument("--kmin", type=float, nargs="+", default=[0.0, 4.0, 8.0, 16.0],
                    help="candidate low-k cutoffs of the closure forcing")
    ap.add_argument("--gain", type=float, nargs="+", default=[1.0],
                    help="candidate gains of the closure forcing (tuned with the cutoff)")
    ap.add_argument("--kclip", type=float, nargs="+", default=[0.0],
                    help="candidate wavenumbers from which the closure forcing is "
                         "backscatter-clipped (closure.clip_forcing; 0 = none), tuned with "
                         "the cutoff and gain"),
    ap.add_argument("--guard", type=float, default=4.0,
                    help="a closure run falls back to the plain solver when its enstrophy "
                         "exceeds this multiple of the window's initial (coarse-grained) one")
    ap.add_argument("--clamp", type=float, nargs="+", default=[0.0],
                    help="candidate output saturations of the network (in units of the "
                         "targets' std; 0 = none), tuned with the cutoff, gain and kclip")
    ap.add_argument("--sign-clip", type=float, nargs="+", default=[0.0],
                    help="candidate strengths in [0, 1] of the network's sign clipping "
                         "(closure.Net.set_sign_clip), tuned with the others")
    ap.add_argument("--ktaper", type=float, nargs="+", default=[0.0],
                    help="candidate shells at which the closure forcing reaches full strength, "
                         "rising linearly from 0 at the cutoff kmin (0 or <= kmin: a hard "
                         "cutoff), tuned with the others")
    ap.add_argument("--period", type=int, nargs="+", default=[1],
                    help="candidate closure refresh periods in steps (forcing held in between)")
    ap.add_argument("--safe-cfl", type=float, nargs=2, default=None, metavar=("ON", "OFF"),
                    help="switch a closure run to the fully sign-clipped network while the "
                         "CFL number of its fixed dt is above ON times the DNS's, back below "
                         "OFF times (default: never)")
    ap.add_argument("--cfl-guard", type=float, default=2.0,
                    help="a closure run falls back when the CFL number of its fixed dt "
                         "exceeds this multiple of the DNS's")
    ap.add_argument("--no-psi-input", action="store_true",
                    help="zero the network's psi input channel (inputs: laplacian of psi, vz)")
    ap.add_argument("--trace", action="store_true",
                    help="log every run's energy and enstrophy at every check")
    ap.add_argument("--params", default=None,
                    help="skip the training: take the network and its normalisation from this "
                         "closure_params.npz (the DNS still runs to reach the same windows)")
    ap.add_argument("--target-kmin", type=float, default=0.0,
                    help="train on the targets' modes |k| >= target_kmin only")
    ap.add_argument("--err-target", type=float, default=None,
                    help="pick the candidate with the largest period whose mean spectrum error on "
                         "the tuning window is at most this (default: the smallest error)")
    ap.add_argument("--realisations", type=int, default=4,
                    help="forcing realisations of each coarse run (averaged)")
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
    out["dns"]["cfl"] = L.cfl_number(ps, vz, p, a.dt)
    log("dns %.4f ms/step, %.3f GPU-s per unit time, CFL number %.3f" % (
        dns_ms, out["dns"]["gpu_s_per_time"], out["dns"]["cfl"]))

    # 2. training window: snapshots
    every = max(1, a.train_steps // a.snapshots)
    if a.params:
        advance(every * a.snapshots)
        z = np.load(a.params)
        params = z["params"]
        scale_in, sy = [float(v) for v in z["scale_in"]], [float(v) for v in z["scale_out"]]
        out["train"] = {"params": a.params}
        for k in ("val_loss", "val_loss_zero_model", "seconds"):
            out["train"][k] = float(z[k]) if k in z else None
    else:
        params, scale_in, sy = train_window(a, advance, ps, vz, out, log)
    np.savez(os.path.join(a.out, "closure_params.npz"), params=params, scale_in=np.array(scale_in),
             scale_out=np.array(sy), nf=nf, nc=nc, **{k: out["train"][k] for k in (
                 "val_loss", "val_loss_zero_model", "seconds") if out["train"].get(k) is not None})

    # 4. tuning window: choose the closure's low-k cutoff on data the final
    # evaluation does not see
    kmax = nc // 3
    dtc = a.dt * nf / nc if a.dt_coarse is None else a.dt_coarse
    seeds = [a.seed + 100 + r for r in range(a.realisations)]
    ps_c0, vz_c0 = L.coarse_grain(ps, nc), L.coarse_grain(vz, nc)
    e0 = S.energy(ps_c0, 1) + S.energy(vz_c0, 0)
    z0 = S.energy(ps_c0, 2) + S.energy(vz_c0, 1)
    # the DNS judge: realisations forced by the coarse runs' draws (the first
    # continues the DNS into the next window)
    dns_tune_runs = dns_realisations(ps, vz, p, a.dt, nc, a.tune_steps, seeds, kf, f0)
    dns_tune = mean_stats(dns_tune_runs)
    log("dns tuning window done (%d realisations)" % len(seeds))
    Tt = a.tune_steps * a.dt
    out["tune"] = {}
    cands = [dict(kmin=km, gain=g, kclip=kc, clamp=cl, sign_clip=sc, period=pe, ktaper=kt)
             for km in a.kmin for g in a.gain for kc in a.kclip for cl in a.clamp
             for sc in a.sign_clip for pe in a.period for kt in a.ktaper]
    key = lambda c: "kmin %g gain %g kclip %g clamp %g sign_clip %g period %d ktaper %g" % (
        c["kmin"], c["gain"], c["kclip"], c["clamp"], c["sign_clip"], c["period"], c["ktaper"])
    taper = lambda c: (L.taper_gain(c["kmin"], c["ktaper"], nc // 2 + 2)
                       if c["ktaper"] > c["kmin"] else None)
    cfl_bound = a.cfl_guard * out["dns"]["cfl"]
    safe = None if a.safe_cfl is None else tuple(v * out["dns"]["cfl"] for v in a.safe_cfl)
    for c in cands:
        stats, steps, finite, fb = coarse_runs(ps_c0, vz_c0, p, Tt, dtc, seeds, kf, f0, params,
                                               scale_in, [c["gain"] * v for v in sy], c["kmin"],
                                               energy_bound=20.0 * e0, kclip=c["kclip"],
                                               enstrophy_bound=a.guard * z0, clamp=c["clamp"],
                                               cfl_bound=cfl_bound, sign_clip=c["sign_clip"],
                                               safe=safe, period=c["period"], gain=taper(c))
        r = {"finite": finite, "fallbacks": fb, "cfl_max": coarse_runs.cfl_max,
             "safe_fraction": coarse_runs.safe_fraction}
        if finite and stats is not None:
            r.update(errors(stats, dns_tune, kmax))
            j = judge(coarse_runs.per_run, dns_tune_runs, kmax)
            r["shells_within_5pct_or_noise"] = j["shells_within_5pct_or_noise"]
            r["shells_within_5pct"] = j["shells_within_5pct"]
        out["tune"][key(c)] = r
        log("tune %s: spectrum err max %s mean %s, shells within 5%% %s (or noise %s), "
            "fallbacks %d, max CFL %.3f, safe %.3f" % (
                key(c), r.get("spectrum_rel_err_max"), r.get("spectrum_rel_err_mean"),
                r.get("shells_within_5pct"), r.get("shells_within_5pct_or_noise"), fb,
                r["cfl_max"], r["safe_fraction"]))
    def err(c):
        r = out["tune"][key(c)]
        return r.get("spectrum_rel_err_mean", 1e300) + (1e300 if r["fallbacks"] else 0.0)

    def agree(c):
        return out["tune"][key(c)].get("shells_within_5pct_or_noise", -1)

    best = min(cands, key=err)
    if a.err_target is not None:
        # the cheapest candidate (fewest network evaluations: largest period)
        # that meets the accuracy target on the tuning window, then the one
        # with the most shells agreeing with the DNS (5% or its noise floor),
        # then the most accurate; the most accurate overall if none meets it
        ok = [c for c in cands if err(c) <= a.err_target]
        if ok:
            best = min(ok, key=lambda c: (-c["period"], -agree(c), err(c)))
    out["selection"] = ("min tuning error" if a.err_target is None else
                        "largest period with tuning error <= %g, then most shells within 5%% "
                        "or the DNS noise floor, then min error" % a.err_target)
    out.update(best)
    kbest, gbest, kcbest, clbest, scbest, pbest, ktbest = (
        best[k] for k in ("kmin", "gain", "kclip", "clamp", "sign_clip", "period", "ktaper"))
    sy = [gbest * v for v in sy]

    # 5. evaluation window: coarse-grained DNS statistics, then the coarse runs
    ps_c0, vz_c0 = L.coarse_grain(ps, nc), L.coarse_grain(vz, nc)
    e0 = S.energy(ps_c0, 1) + S.energy(vz_c0, 0)
    z0 = S.energy(ps_c0, 2) + S.energy(vz_c0, 1)
    dns_runs = dns_realisations(ps, vz, p, a.dt, nc, a.eval_steps, seeds, kf, f0)
    dns_stats = mean_stats(dns_runs)
    T = a.eval_steps * a.dt
    log("dns evaluation window done (t = %.3f); closure %s" % (T, key(best)))
    del ps, vz
    torch.cuda.empty_cache()
    for name, prm in (("plain", None), ("closure", params)):
        stats, steps, finite, fb = coarse_runs(ps_c0, vz_c0, p, T, dtc, seeds, kf, f0, prm, scale_in,
                                               sy, kbest, energy_bound=20.0 * e0, kclip=kcbest,
                                               enstrophy_bound=a.guard * z0,
                                               trace=log if a.trace and prm is not None else None,
                                               clamp=clbest, cfl_bound=cfl_bound,
                                               sign_clip=scbest, safe=safe, period=pbest,
                                               gain=taper(best))
        net = L.Net(1, nc, nc, dev, params=params) if prm is not None else None
        if net is not None:
            net.set_forcing_gain(taper(best))
        sec = timed_steps(ps_c0, vz_c0, p, dtc, Forcing(kf, f0, a.seed + 11, nc), 500, net, scale_in,
                          sy, kbest, kcbest, clbest, scbest, pbest)
        r = {"finite": finite, "steps": steps, "dt": dtc, "sim_time": steps * dtc,
             "realisations": len(seeds), "ms_per_step": sec * 1e3, "gpu_s_per_time": sec / dtc,
             "fallbacks": fb, "cfl_max": coarse_runs.cfl_max if prm is not None else None,
             "safe_fraction": coarse_runs.safe_fraction if prm is not None else None}
        if finite and stats is not None:
            r.update(errors(stats, dns_stats, kmax))
            r["spectrum"] = [float(x) for x in (stats[0] + stats[1])[:kmax]]
            r["judge"] = judge(coarse_runs.per_run, dns_runs, kmax)
        out["coarse"][name] = r
        log("%s: finite %s, %d steps, t %.3f, %.3f GPU-s/time, spectrum err max %s mean %s "
            "k<=40 max %s, flux err %s" % (
                name, finite, steps, r["sim_time"], r["gpu_s_per_time"],
                r.get("spectrum_rel_err_max"), r.get("spectrum_rel_err_mean"),
                r.get("spectrum_rel_err_k_le_40_max"), r.get("flux_err")))
    out["dns"]["spectrum"] = [float(x) for x in (dns_stats[0] + dns_stats[1])[:kmax]]
    c = out["coarse"]["closure"]
    out["speedup_gpu_s_per_time"] = out["dns"]["gpu_s_per_time"] / c["gpu_s_per_time"]
    out["wall_seconds"] = time.perf_counter() - t0
    with open(os.path.join(a.out, "closure_study.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez(os.path.join(a.out, "closure_params.npz"), params=params, scale_in=np.array(scale_in),
             scale_out=np.array(sy), kmin=kbest, gain=gbest, kclip=kcbest,
             out_clamp=clbest, sign_clip=scbest, period=pbest, ktaper=ktbest, nf=nf, nc=nc,
             **({} if safe is None else {"safe_cfl": np.array(safe)}), **{k: out["train"][k] for k in (
                 "val_loss", "val_loss_zero_model", "seconds") if out["train"].get(k) is not None})
    cmd = "python -m rot2d3c_amd.closure_study " + " ".join(
        sys.argv[1:] if argv is None else list(argv))
    with open(os.path.join(a.out, "closure_eval.json"), "w") as f:
        json.dump(compact_record(out, cmd), f, indent=1)
    log("speedup %.1fx; wrote %s" % (out["speedup_gpu_s_per_time"], a.out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
