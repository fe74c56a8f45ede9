#!/usr/bin/env python3
"""Benchmark: 2048^2 FP64 Rot2D3C RK4 steps on one MI355X.

Workload (the single-GPU run the headline metric is quoted on): n = 2048,
fixed dt = 2e-4, nu = nuv = 5e-5, hnu = hnuv = 0.5 (nn = nnv = 1, mm = mmv = 2),
Omega = 5, random forcing (iflow = 3) on the shell kf = 40 for psi and vz with
injection rate 1.0, forcing seed 0; seeded random-phase initial condition.
One step = 4 x (row + fused column) kernels; the steps of the timed region
run as one r2d3c_run call, so the column transforms of each step's state come
from the previous step's last kernel (one column-inverse kernel per call).
Inputs resident in HBM.

Prints one JSON line.  `value` is Mpts*RK4-steps/s = n^2*steps/s/1e6.  The
`roofline` object describes the kernel class with the largest share of the
step time, timed live with HIP events stamped at kernel start/end
(hipExtLaunchKernelGGL) over a second timed region of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E

from rot2d3c_amd.traffic import algorithmic_bytes  # noqa: E402


def cpu_baseline(n, steps, p, kf, f0, seed, ps0, vz0, dt, workers=None):
    """The numpy oracle on the same workload for `steps` steps, forcing drawn
    from the same seed; its 2D FFTs on `workers` threads (None: numpy, one
    core; the pointwise work stays single-threaded numpy).  Returns the
    elapsed seconds."""
    from oracle import spectral as O
    op = O.Params(p["nu"], p["hnu"], p["nn"], p["mm"], p["nuv"], p["hnuv"], p["nnv"],
                  p["mmv"], p["omega"])
    O.set_fft_workers(workers)
    try:
        t0 = time.perf_counter()
        O.run_random_forcing(ps0.copy(), vz0.copy(), steps, 0.0, op, kf, kf, f0, f0, seed,
                             fixed_dt=dt)
        return time.perf_counter() - t0
    finally:
        O.set_fft_workers(None)


def cpu_baseline_compiled(n, steps, p, kf, f0, seed, ps0, vz0, dt, threads):
    """The compiled C++/OpenMP restatement of the step (oracle/cpu_step.cpp:
    the reference's 40 full 2D transforms per fixed-dt step, its own FFT) on
    `threads` host threads, same workload and forcing draws.  Returns the
    elapsed seconds."""
    from oracle import cpu_step as C
    from oracle import spectral as O
    if not os.path.exists(C.LIB):
        C.build()
    C.threads(threads)
    op = O.Params(p["nu"], p["hnu"], p["nn"], p["mm"], p["nuv"], p["hnuv"], p["nnv"],
                  p["mmv"], p["omega"])
    s0, s1 = seed, seed + 1
    _, _, _, s0 = O.rand_force_draw(kf, s0)  # the pre-loop pair advances both seeds
    _, _, _, s1 = O.rand_force_draw(kf, s1)
    draws = []
    for _ in range(steps):
        kx, ky, ph, s0 = O.rand_force_draw(kf, s0)
        a = (kx, ky, ph, f0, kf)
        kx, ky, ph, s1 = O.rand_force_draw(kf, s1)
        draws.append((a, (kx, ky, ph, f0, 1.0)))
    a, b = ps0.copy(), vz0.copy()
    t0 = time.perf_counter()
    C.run(a, b, op, steps, dt, draws)
    return time.perf_counter() - t0


def closure_bench(dns_ms_per_step, dns_dt, dev, steps=500, warmup=20, judge_window=10000,
                  judge_realisations=2):
    """GPU-seconds per unit simulated time of the 256^2 run with the learned
    closure (the committed trained network, rot2d3c_amd/data/closure_256.npz;
    one-run ensemble, closure inside the step graphs with its committed
    settings, fixed dt = 8 x the DNS dt: the DNS's CFL number; chunks of 50
    steps, each followed by the run-time stabiliser's CFL check) against the
    2048^2 DNS timed above.  The spectrum and flux errors are measured here
    (rot2d3c_amd.closure_study.live_evaluation): the 2048^2 DNS is spun up
    from its seeded start, `judge_realisations` DNS runs over `judge_window`
    steps are coarse-grained and averaged, and the closure and plain 256^2
    runs from the coarse-grained start, forced by the same draws, are judged
    against them for k <= 85 (a shell agrees within 5% or within the DNS
    noise floor measured from the realisations' spread)."""
    import torch
    from rot2d3c_amd import closure as L, closure_study as CS, ensemble as E, ic, spectral as S
    path = os.path.join(ROOT, "rot2d3c_amd", "data", "closure_256.npz")
    z = np.load(path)
    cs = CS.closure_settings(z)
    n, dt = int(z["nc"]), dns_dt * int(z["nf"]) / int(z["nc"])
    ps, vz = ic.random_phase_ic(n, seed=7)
    p = S.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    # the default schedule, as the closure study that evaluated the network
    ens = E.Ensemble([E.Member(torch.from_numpy(ps).to(dev), torch.from_numpy(vz).to(dev), p, dt)],
                     schedules=None)
    net = L.Net(1, n, n, dev, params=z["params"])
    if cs["ktaper"] > cs["kmin"]:
        net.set_forcing_gain(L.taper_gain(cs["kmin"], cs["ktaper"], n // 2 + 2))
    if "out_clamp" in z:
        net.set_output_clamp(float(z["out_clamp"]))
    if "sign_clip" in z:
        net.set_sign_clip(float(z["sign_clip"]))
    ens.set_closure(0, net, z["scale_in"], z["scale_out"], float(z["kmin"]),
                    float(z["kclip"]) if "kclip" in z else 0.0)
    if "period" in z:
        ens.set_closure_period(int(z["period"]))
    spec = [dict(kf=40.0, f0=1.0, seed=3)]
    warm = ens.forcing_array(E.forcing_table(spec, warmup))
    # chunks of 50 steps with the run-time stabiliser's CFL check after each,
    # as the evaluated closure runs
    chunk = 50
    steps = max(chunk, steps // chunk * chunk)
    full = E.forcing_table(spec, steps)
    tables = [ens.forcing_array(full[i:i + chunk]) for i in range(0, steps, chunk)]
    sm = L.SafeMode(*[float(v) for v in z["safe_cfl"]], float(z["sign_clip"])) \
        if "safe_cfl" in z else None
    ens.run(warmup, warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for tab in tables:
        ens.run(chunk, tab)
        if sm is not None and sm.check(0, L.cfl_number(ens.members[0].ps, ens.members[0].vz, p, dt)):
            sm.apply(net, sm.active[0])
            ens.set_closure(0, net, z["scale_in"], z["scale_out"], float(z["kmin"]),
                            float(z["kclip"]) if "kclip" in z else 0.0)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    finite = bool(torch.isfinite(torch.view_as_real(ens.members[0].ps)).all().item()
                  and torch.isfinite(torch.view_as_real(ens.members[0].vz)).all().item())
    ens.close()
    net.close()
    les = ms * 1e-3 / dt
    dns = dns_ms_per_step * 1e-3 / dns_dt
    out = {"workload": "256^2 run with the learned closure (BF16 MFMA CNN) vs 2048^2 DNS",
           "speedup_vs_dns": round(dns / les, 2), "les_ms_per_step": round(ms, 4), "les_dt": dt,
           "les_gpu_s_per_time": round(les, 4), "dns_gpu_s_per_time": round(dns, 4),
           "timed_run_finite": finite}
    if judge_window <= 0:
        return out
    # accuracy, measured here against a DNS window generated here
    jd = CS.live_evaluation(path, dev, window=judge_window, realisations=judge_realisations)
    c, pl = jd["closure"], jd["plain"]
    r4 = lambda x: None if x is None else round(x, 4)
    return {**out,
            "spectrum_rel_err_mean": r4(c.get("spectrum_rel_err_mean")),
            "spectrum_rel_err_max": r4(c.get("spectrum_rel_err_max")),
            "flux_err": r4(c.get("flux_err")),
            "shells_within_5pct": "%s/%s" % (c.get("shells_within_5pct"), c.get("shells")),
            "shells_within_5pct_or_dns_noise": "%s/%s" % (c.get("shells_within_5pct_or_noise"),
                                                         c.get("shells")),
            "shells_outside": c.get("shells_outside"),
            "plain_spectrum_rel_err_mean": r4(pl.get("spectrum_rel_err_mean")),
            "plain_flux_err": r4(pl.get("flux_err")),
            "closure_finite": c.get("finite"), "fallbacks": c.get("fallbacks"),
            "stabiliser_time_fraction": c.get("safe_fraction"),
            "errors_from": "measured in this run: 2048^2 DNS spun up %d steps, %d realisations of "
                           "%.1f time units forced by the coarse runs' draws, coarse-grained; "
                           "closure and plain 256^2 runs (as many realisations) from the "
                           "coarse-grained start; k <= 85" % (
                               jd["spinup_steps"], judge_realisations, jd["window_time"]),
            "settings": jd["settings"]}


def large_bench(steps, dev, warmup=2):
    """The large single run: 16384^2 FP64, fixed dt = 2e-5, nu = nuv = 5e-6,
    hnu = hnuv = 1.0, Omega = 10, forcing kf = 256 (psi and vz, injection
    1.0), random-phase start.  At this size the step runs from 2D transforms
    of the dealiased columns (two 8192-point halves per transform) with the
    fused bracket pass (both Jacobians and their row transforms in one pass,
    the update inside the second column transform); the stage-by-stage form
    (six real derivative fields, separate update) is timed beside it on
    min(steps, 20) further steps."""
    import torch
    from rot2d3c_amd import _lib, ic, spectral as S
    n, dt = 16384, 2e-5
    ps0, vz0 = ic.random_phase_ic(n, seed=1234)
    ps = torch.from_numpy(ps0).to(dev)
    vz = torch.from_numpy(vz0).to(dev)
    del ps0, vz0
    p = S.Params(5e-6, 1.0, 1, 2, 5e-6, 1.0, 1, 2, 10.0)
    nref = min(steps, 20)
    sched = ic.forcing_schedule(256.0, 1.0, 1.0, 0, warmup + steps + nref, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    plan = S.Plan.get(n, dev).handle
    _lib.check(_lib.lib().r2d3c_plan_set_bracket_pass(plan, 1), "set_bracket_pass")
    S.rk4_run(ps, vz, p, draws[:warmup], warmup, dt=dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.rk4_run(ps, vz, p, draws[warmup:warmup + steps], steps, dt=dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    finite = bool(torch.isfinite(torch.view_as_real(ps)).all())
    ms = wall * 1e3 / steps
    ms_ref = None
    if nref > 0:  # the stage-by-stage form on the steps that follow
        _lib.check(_lib.lib().r2d3c_plan_set_bracket_pass(plan, 0), "set_bracket_pass")
        t0 = time.perf_counter()
        S.rk4_run(ps, vz, p, draws[warmup + steps:], nref, dt=dt)
        torch.cuda.synchronize()
        ms_ref = round((time.perf_counter() - t0) * 1e3 / nref, 3)
        _lib.check(_lib.lib().r2d3c_plan_set_bracket_pass(plan, 1), "set_bracket_pass")
        finite = finite and bool(torch.isfinite(torch.view_as_real(ps)).all())
    b76 = 76 * 16 * n * (n // 2 + 1)
    S.Plan.release(n, dev)
    del ps, vz
    torch.cuda.empty_cache()
    return {"workload": "16384^2 FP64, dt=2e-5, nu=5e-6, nu+=1.0, Omega=10, kf=256, "
                        "step with the fused bracket pass",
            "kernels": "profiles/step16384_kernels.md",
            "steps": steps, "ms_per_step": round(ms, 3),
            "stage_by_stage_ms_per_step": ms_ref, "stage_by_stage_steps": nref,
            "value": round(n * n * steps / wall / 1e6, 2), "unit": "Mpts·RK4-steps/s",
            "model_76S_frac_of_peak": round(b76 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "finite": finite}


def copy_probe(dev, nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth: the library's streaming copy kernel
    (r2d3c_copy_probe) on nbytes, best of `reps` calls timed with events on
    the stream it runs on; read + write = 2 nbytes per call."""
    import torch
    from rot2d3c_amd import _lib, spectral as S
    src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    lib = _lib.lib()
    stream = S._stream(dev)
    for _ in range(3):
        _lib.check(lib.r2d3c_copy_probe(dst.data_ptr(), src.data_ptr(), nbytes, stream), "copy_probe")
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        _lib.check(lib.r2d3c_copy_probe(dst.data_ptr(), src.data_ptr(), nbytes, stream), "copy_probe")
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    ok = bool(torch.equal(dst, src))
    del src, dst
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("copy probe produced a wrong copy")
    return 2 * nbytes / (best * 1e-3) / 1e9


def ensemble_bench(steps, dev, tuned, warmup=10):
    """The batched workload: 24 x 256^2 + 12 x 512^2 + 4 x 1024^2 independent
    runs (per-run seed, Omega, nu, forcing shell, fixed dt; see
    rot2d3c_amd.ensemble.mixed_ensemble_spec) advanced together, one captured
    HIP graph per step.  Returns the throughput summed over the runs."""
    import torch
    from rot2d3c_amd import ensemble as E
    spec = E.mixed_ensemble_spec()
    ens = E.build(spec, dev, schedules=None)
    if tuned:  # schedules timed inside the ensemble's own step graph
        ens.tune()
    table = E.forcing_table(spec, warmup + steps)
    ens.run(warmup, table[:warmup])
    # the draws as the library's forcing array before the timed region (the
    # conversion of 20 000 Python draws is harness work, ~2.4 us each); the
    # library's own table set-up and upload stay inside
    fo = ens.forcing_array(table[warmup:])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps, fo)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(torch.view_as_real(m.ps)).all()) for m in ens.members)
    schedules = {str(n): list(v) for n, v in ens.schedules.items()}
    row_split = {str(n): v for n, v in ens.row_split.items()}
    form = ens.stage_launch
    names = {0: "per-size kernels", 1: "one persistent launch per stage", 2: "one launch per phase"}
    forms_ms = {names[form]: round(1e3 * wall / steps, 4)}
    # the other stage launch forms on the steps that follow (their own
    # schedules), for comparison
    nalt = min(steps, 100)
    talt = E.forcing_table([dict(r, seed=r["seed"] + 1000) for r in spec], nalt + 2)
    foalt = ens.forcing_array(talt[2:])
    for f in (0, 1, 2):
        if f == form:
            continue
        ens.set_stage_launch(f)
        ens.run(2, talt[:2])
        ens.sync()
        t0 = time.perf_counter()
        ens.run(nalt, foalt)
        ens.sync()
        forms_ms[names[f] + (" (their schedules)" if f == 0 else "")] = round(
            1e3 * (time.perf_counter() - t0) / nalt, 4)
    ens.close()
    pts = E.points(spec)
    alg = sum(algorithmic_bytes(r["n"], steps)[1] for r in spec)
    b76 = sum(algorithmic_bytes(r["n"], steps)[2]["B76S"] for r in spec)
    ms = 1e3 * wall / steps
    return {"workload": "24 x 256^2 + 12 x 512^2 + 4 x 1024^2 FP64 runs, own seed, Omega~U[0,10], "
                        "nu=1e-3*256/n, kf=n/32, fixed dt=1e-3*256/n",
            "runs": len(spec), "points": pts, "steps": steps, "ms_per_step": round(ms, 4),
            "value": round(pts * steps / wall / 1e6, 2), "unit": "Mpts·RK4-steps/s",
            "algorithmic_bytes_per_step": alg,
            "frac_of_peak": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "model_76S_frac_of_peak": round(b76 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "finite": finite, "schedules": schedules,
            "size_classes": {0: "one after another", 1: "concurrent graph branches",
                             2: "each on a stream of its own"}[ens.concurrent],
            "row_split": row_split, "stage_launch": names[form],
            "stage_launch_forms_ms_per_step": forms_ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="steps of the oracle timed on the host (0 disables)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default row schedule instead of the cached autotuned one")
    ap.add_argument("--large-steps", type=int, default=200,
                    help="steps of the 16384^2 run (0 disables)")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the closure-vs-DNS timing")
    ap.add_argument("--closure-judge-window", type=int, default=10000,
                    help="DNS steps of each realisation of the closure's live judge "
                         "(0: time the closure run only, no judge)")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="steps of the batched ensemble (24 x 256^2 + 12 x 512^2 + 4 x 1024^2); "
                         "0 disables")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 16384^2 section and print its object (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if args.large_only:
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        print(json.dumps({"large": large_bench(max(args.large_steps, 1), dev)}), flush=True)
        return
    from rot2d3c_amd import _lib
    from rot2d3c_amd import ic, spectral as S

    n = args.n
    dt = 2e-4
    p = dict(nu=5e-5, hnu=0.5, nn=1, mm=2, nuv=5e-5, hnuv=0.5, nnv=1, mmv=2, omega=5.0)
    kf, f0, seed = 40.0, 1.0, 0
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ps0, vz0 = ic.random_phase_ic(n, seed=1234)
    total = args.warmup + args.steps + 6 * min(args.steps, 100)
    sched = ic.forcing_schedule(kf, f0, f0, seed, total, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    sp = S.Params(**p)
    ps = torch.from_numpy(ps0).to(dev)
    vz = torch.from_numpy(vz0).to(dev)
    plan = S.Plan.get(n, dev)
    from rot2d3c_amd import autotune
    sch = autotune.DEFAULT
    if not args.no_autotune:
        sch = autotune.apply(plan.handle, n, dev)

    def run(k0, k):
        # r2d3c_run: k chained steps (bitwise equal to k single steps)
        S.rk4_run(ps, vz, sp, draws[k0:k0 + k], dt=dt)

    run(0, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.warmup, args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n * n * args.steps / wall / 1e6

    # second timed region: per-kernel durations from start/stop-stamped events
    ktimed = min(args.steps, 100)
    lib = _lib.lib()
    _lib.check(lib.r2d3c_timing_enable(plan.handle, 1), "timing_enable")
    run(args.warmup + args.steps, ktimed)
    ms = (_lib.ctypes.c_double * _lib.R2D3C_NKERNELS)()
    cnt = (_lib.ctypes.c_longlong * _lib.R2D3C_NKERNELS)()
    _lib.check(lib.r2d3c_timing_read(plan.handle, ms, cnt), "timing_read")
    _lib.check(lib.r2d3c_timing_enable(plan.handle, 0), "timing_disable")
    torch.cuda.synchronize()
    # run-to-run spread: five more timed runs of ktimed steps each
    reps = []
    for r in range(5):
        k0 = args.warmup + args.steps + (1 + r) * ktimed
        t0 = time.perf_counter()
        run(k0, ktimed)
        torch.cuda.synchronize()
        reps.append(1e3 * (time.perf_counter() - t0) / ktimed)
    repeats = {"runs": 5, "steps_each": ktimed, "ms_per_step_median": round(float(np.median(reps)), 4),
               "ms_per_step_min": round(min(reps), 4), "ms_per_step_max": round(max(reps), 4)}
    finite = bool(torch.isfinite(torch.view_as_real(ps)).all() and torch.isfinite(torch.view_as_real(vz)).all())

    per, step_bytes, geo = algorithmic_bytes(n, args.steps)
    per, _, _ = algorithmic_bytes(n, ktimed)  # the event-timed region is one call of ktimed steps
    kern = {}
    for k, name in enumerate(_lib.KERNEL_NAMES):
        if cnt[k]:
            avg_ms = ms[k] / cnt[k]
            kern[name] = {"avg_us": 1e3 * avg_ms, "launches_per_step": cnt[k] / ktimed,
                          "share": ms[k] / sum(ms), "bytes_per_launch": per[name],
                          "GBps": per[name] / (avg_ms * 1e-3) / 1e9}
    dom = max(kern, key=lambda k: kern[k]["share"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_step2048.json")
    if os.path.exists(tfile) and n == 2048:
        with open(tfile) as f:
            tj = json.load(f)
        if dom in tj.get("kernels", {}):
            traffic = tj["kernels"][dom]["hbm_bytes_per_launch"]
    achieved = kern[dom]["GBps"]
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "kernel": dom,
            "bytes_per_launch": per[dom], "avg_launch_us": round(kern[dom]["avg_us"], 2)}

    copy_gbps = copy_probe(dev)
    achievable = {"copy_GBps": round(copy_gbps, 1),
                  "what": "r2d3c_copy_probe: 1 GiB device-to-device streaming copy, "
                          "2 x 1 GiB of HBM traffic per call, best of 20 event-timed calls",
                  "dominant_kernel_frac_of_copy": round(achieved / copy_gbps, 4),
                  "step_76S_frac_of_copy": round(geo["B76S"] / (ms_per_step * 1e-3) / 1e9
                                                 / copy_gbps, 4)}

    cpu = None
    if args.cpu_steps > 0:
        # host cores this process may use (the GPU's share of the host)
        cores = min(16, int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1,
                    len(os.sched_getaffinity(0)))
        k = args.cpu_steps
        eln = cpu_baseline_compiled(n, k, p, kf, f0, seed, ps0, vz0, dt, cores)
        el1 = cpu_baseline_compiled(n, 1, p, kf, f0, seed, ps0, vz0, dt, 1)
        elp = cpu_baseline(n, 1, p, kf, f0, seed, ps0, vz0, dt, workers=cores)
        mp = lambda steps, sec: round(n * n * steps / sec / 1e6, 4)
        cpu = {"value": mp(k, eln), "unit": "Mpts·RK4-steps/s", "cores": cores, "kind": "port",
               "sample": "%d steps of the same %d^2 workload through the compiled C++/OpenMP "
                         "restatement (oracle/cpu_step.cpp: 40 full 2D FFTs per step, own "
                         "radix-4 FFT) on %d threads: %.1f s; 1 step on one thread: %.1f s; "
                         "1 step of the numpy oracle (FFTs threaded over %d cores): %.1f s"
                         % (k, n, cores, eln, el1, cores, elp),
               "value_one_core": mp(1, el1), "value_numpy_oracle": mp(1, elp)}

    clo = None
    if not args.no_closure:
        clo = closure_bench(ms_per_step, dt, dev, judge_window=args.closure_judge_window)

    big = None
    if args.large_steps > 0:
        big = large_bench(args.large_steps, dev)

    ens = None
    if args.ensemble_steps > 0:
        del ps, vz
        torch.cuda.empty_cache()
        ens = ensemble_bench(args.ensemble_steps, dev, not args.no_autotune)

    out = {
        "metric": "Mpts·RK4-steps/s at 2048² FP64 (%HBM peak); closure-vs-DNS speedup at 5% err",
        "value": round(value, 2),
        "unit": "Mpts·RK4-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase initial condition, seeded random forcing)",
        "config": {"workload": "Rot2D3C RK4 step, %d^2 FP64, dt=2e-4 fixed, nu=5e-5, nu+=0.5, "
                               "Omega=5, forcing kf=40 (psi and vz), injection 1.0" % n,
                   "grid": n, "dt": dt, "parallelism": "single GPU",
                   "schedule": {"row_rows": sch[0], "row_blocks_per_cu": sch[1],
                                "column_radix_form": sch[2]}},
        "roofline": roof,
        "achievable_hbm": achievable,
        "repeats": repeats,
        "cpu_baseline": cpu,
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes,
                     "achieved_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                     "frac_of_peak": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                     "model_76S_bytes_per_step": geo["B76S"],
                     "model_76S_frac_of_peak": round(geo["B76S"] / (ms_per_step * 1e-3) / 1e9
                                                     / HBM_PEAK_GBPS, 4)},
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
        "finite": finite,
        "ensemble": ens,
        "closure": clo,
        "large": big,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
