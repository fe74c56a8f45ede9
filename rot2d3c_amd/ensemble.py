"""Ensembles of independent runs on one GPU (no reference counterpart: the
reference runs one simulation per MPI job).

Runs of one grid size advance together: every kernel of the step covers all
of them in one launch (the run is blockIdx.y), and each size class runs on a
stream of its own, so that the classes overlap on the device (or the classes
follow each other, set_concurrent), from HIP graphs captured once and
replayed: one of 8 steps for the steady steps (a multiple of the closure
period) and one-step graphs for the steps around them.  Two other launch forms cover every size at once
(set_stage_launch): one persistent launch per RK4 stage, or one launch per
phase.  The forcing of every step is drawn on the host before the steps start
(rand_force's wavevectors and phases do not depend on the flow) and uploaded
as one table, so the steps need no host round trip.  Each run equals the same
run advanced alone by rk4_run with its fixed dt, bit for bit.
"""
import ctypes
import time
import math

import numpy as np
import torch

from . import _lib, autotune
from . import spectral as S
from ._lib import check


class Member:
    """One run: grid size n, parameters (spectral.Params), fixed dt and its
    device state ps, vz (complex128, (n/2+1, n), advanced in place)."""

    def __init__(self, ps, vz, params, dt):
        self.n = S._spec_n(ps)
        if S._spec_n(vz) != self.n or ps.device != vz.device:
            raise ValueError("psi and vz of a run must match")
        self.ps, self.vz, self.params, self.dt = ps, vz, params, float(dt)

    def c(self):
        return _lib.Member(self.n, self.params.c(), self.dt, S._ptr(self.ps), S._ptr(self.vz))


class Ensemble:
    def __init__(self, members, schedules="auto"):
        """members: list of Member on one device.  schedules: "auto" (cached
        autotuned schedule per size), None (defaults) or {n: (rows, bpc)} /
        {n: (rows, bpc, column_form)} (autotune module: schedule meanings)."""
        if not members:
            raise ValueError("empty ensemble")
        self.members = list(members)
        self.device = self.members[0].ps.device
        arr = (_lib.Member * len(self.members))(*[m.c() for m in self.members])
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_create(ctypes.byref(h), arr, len(arr)),
                  "r2d3c_ensemble_create")
        self.handle = h
        self.sizes = sorted({m.n for m in self.members})
        self.concurrent = 2
        self.stage_launch = 0
        self.row_split = {n: -1 for n in self.sizes}
        self.schedules = {n: autotune.DEFAULT for n in self.sizes}
        for n in self.sizes:
            if schedules == "auto":
                sch = autotune.schedule(n, self.device)
            elif schedules:
                sch = schedules.get(n, autotune.DEFAULT)
            else:
                continue
            check(self._set_schedule(n, sch), "r2d3c_ensemble_set_schedule")

    def _set_schedule(self, n, sch):
        """Set size n's schedule (rows, blocks per CU[, column form]); 0 or
        the library's error code (the schedule is then unchanged)."""
        rw, bpc = int(sch[0]), int(sch[1])
        cs = int(sch[2]) if len(sch) > 2 else 0
        old = self.schedules[n]
        rc = _lib.lib().r2d3c_ensemble_set_schedule(self.handle, n, rw, bpc)
        if rc == 0:
            rc = _lib.lib().r2d3c_ensemble_set_column_radix(self.handle, n, cs)
            if rc != 0:
                _lib.lib().r2d3c_ensemble_set_schedule(self.handle, n, old[0], old[1])
        if rc == 0:
            self.schedules[n] = (rw, bpc, cs)
        return rc

    def tune(self, steps=20, trials=5, log=None):
        """Choose the order of the size classes (one after another or
        concurrent), every size's schedule and row form, and the stage launch
        form by timing the whole ensemble step graph (a solo plan's fastest
        schedule need not be the fastest there): the order first, then one
        coordinate-descent sweep over the sizes, largest first, each trying
        every valid schedule and then the other row form with the other sizes
        fixed, then the all-size stage launches (best of `trials` timings of
        `steps` unforced steps).
        The runs' states are restored afterwards, so tuning leaves the
        ensemble's trajectories as they were.  Returns the schedules and the
        final milliseconds per ensemble step."""
        saved = [(m.ps.clone(), m.vz.clone()) for m in self.members]

        def timed():
            self.run(2)  # recapture and warm up
            torch.cuda.synchronize(self.device)
            best = float("inf")
            for _ in range(trials):
                t0 = time.perf_counter()
                self.run(steps)
                torch.cuda.synchronize(self.device)
                best = min(best, (time.perf_counter() - t0) * 1e3 / steps)
            if log:
                log("tune: %s concurrent=%d row_split=%s stage_launch=%d: %.4f ms" % (
                    dict(self.schedules), self.concurrent, dict(self.row_split),
                    self.stage_launch, best))
            return best

        best_ms = timed()
        # the order of the size classes in a step: one after another (the
        # default), as concurrent graph branches, or each on its own stream
        if len(self.sizes) > 1:
            keep = self.concurrent
            for mode in (0, 1, 2):
                if mode == keep:
                    continue
                self.set_concurrent(mode)
                ms = timed()
                if ms < best_ms:
                    best_ms, keep = ms, mode
            self.set_concurrent(keep)
        for n in sorted(self.sizes, reverse=True):
            # every schedule, and with the non-persistent row kernel at
            # n <= 512 both row forms (the parallel field pairs, automatic
            # there and made for latency-bound launches, and the sequential
            # one, which a large batched launch may prefer)
            for cand in autotune.CANDIDATES:
                if cand[2] and n < 512:
                    continue
                for split in ((-1, 0) if n <= 512 and cand[1] == 0 else (-1,)):
                    if tuple(cand) == tuple(self.schedules[n]) and split == self.row_split[n]:
                        continue
                    keep = (self.schedules[n], self.row_split[n])
                    if self._set_schedule(n, cand) != 0:
                        continue
                    self.set_row_split(n, split)
                    ms = timed()
                    if ms < best_ms:
                        best_ms = ms
                    else:
                        check(self._set_schedule(n, keep[0]), "r2d3c_ensemble_set_schedule")
                        self.set_row_split(n, keep[1])
        # the all-size stage launches (one persistent launch per stage, one
        # launch per phase), where every size fits them
        if all(64 <= n <= 1024 for n in self.sizes) and len(self.sizes) <= 5 and \
                not getattr(self, "_nets", None):
            keep = (dict(self.schedules), dict(self.row_split))
            best_form = 0
            for form in (1, 2):
                self.set_stage_launch(form)
                ms = timed()
                if ms < best_ms:
                    best_ms, best_form = ms, form
            self.set_stage_launch(best_form)
            if best_form == 0:
                for n in self.sizes:
                    check(self._set_schedule(n, keep[0][n]), "r2d3c_ensemble_set_schedule")
                    self.set_row_split(n, keep[1][n])
        for m, (ps, vz) in zip(self.members, saved):
            m.ps.copy_(ps)
            m.vz.copy_(vz)
        torch.cuda.synchronize(self.device)
        return dict(self.schedules), best_ms

    def set_stage_launch(self, form):
        """How a step's stages are launched: 0 the per-size kernels (default),
        1 one persistent launch per RK4 stage for every size (a grid barrier
        between the row and column phases), 2 one launch per phase for every
        size.  Forms 1 and 2 take sizes 64..1024 and set every size's
        schedule to theirs (sequential row kernel, radix-8 Stockham columns:
        autotune schedule (2, 0, 1), (2, 0, 0) below 512, row split off)."""
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_set_stage_launch(self.handle, int(form)),
                  "r2d3c_ensemble_set_stage_launch")
        self.stage_launch = int(form)
        if form:  # (the column form is moot below 512)
            self.schedules = {n: (2, 0, 1 if n >= 512 else 0) for n in self.sizes}
            self.row_split = {n: 0 for n in self.sizes}

    def set_row_split(self, n, mode):
        """Row kernel form of size n (-1 automatic: parallel field pairs for
        n <= 512; 0 sequential; 1 parallel)."""
        check(_lib.lib().r2d3c_ensemble_set_row_split(self.handle, int(n), int(mode)),
              "r2d3c_ensemble_set_row_split")
        self.row_split[n] = int(mode)

    def persistent_grid(self):
        """Blocks of the persistent launch (form 1), else 0."""
        form, grid = ctypes.c_int(0), ctypes.c_int(0)
        check(_lib.lib().r2d3c_ensemble_get_stage_launch(self.handle, ctypes.byref(form),
                                                         ctypes.byref(grid)),
              "r2d3c_ensemble_get_stage_launch")
        return int(grid.value)

    def sync(self):
        """Wait for the ensemble's work; raises if a grid barrier of the
        persistent launch timed out (that run call's states are invalid)."""
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_sync(self.handle), "r2d3c_ensemble_sync")

    def set_concurrent(self, concurrent):
        """Size classes of a step one after another (0), as concurrent
        branches of the step graph (1), or each on a stream of its own for a
        whole run call (2, the default; ensembles with a closure run and the
        all-size stage launches take 0); results are identical."""
        check(_lib.lib().r2d3c_ensemble_set_concurrent(self.handle, int(concurrent)),
              "r2d3c_ensemble_set_concurrent")
        self.concurrent = int(concurrent)

    def set_closure(self, run, net=None, scale_in=None, scale_out=None, kmin=0.0, kclip=0.0):
        """Make run `run` a closure run with network `net` (closure.Net for one
        n x n image, used by this run only; its forcing acts on |k| >= kmin,
        backscatter-clipped on |k| >= kclip if kclip > 0), or the plain solver
        (net=None)."""
        self._nets = getattr(self, "_nets", {})
        d = lambda v: (ctypes.c_double * len(v))(*[float(x) for x in v])
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_set_closure(
                self.handle, int(run), net.handle if net is not None else None,
                d(scale_in) if net is not None else None, d(scale_out) if net is not None else None,
                float(kmin), float(kclip)),
                "r2d3c_ensemble_set_closure")
        self._nets[run] = net  # keep the network alive while the ensemble uses it

    def set_closure_period(self, period):
        """Evaluate the closure runs' forcing on every period-th step of each
        run() call (its first step, then steps period, 2 period, ...) and
        hold it in between."""
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_set_closure_period(self.handle, int(period)),
                  "r2d3c_ensemble_set_closure_period")

    def forcing_array(self, forcing):
        """The C forcing table of `forcing` (see run), reusable across calls."""
        n_steps = len(forcing)
        if any(len(f) != len(self.members) for f in forcing):
            raise ValueError("forcing must be lists of one (fp, fz) pair per run")
        fo = (_lib.Forcing * (n_steps * len(self.members)))()
        for st, row in enumerate(forcing):
            for m, pair in enumerate(row):
                fp, fz = pair if pair is not None else (None, None)
                fo[st * len(self.members) + m] = S._forcing_c(fp, fz)
        return fo

    def run(self, n_steps, forcing=None):
        """Advance every run n_steps steps on the current stream.  forcing:
        None, a list over steps of a list over runs of (fp_draw, fz_draw)
        (spectral.ForceDraw or None each), or an array from forcing_array."""
        fo = None
        if forcing is not None:
            if isinstance(forcing, ctypes.Array):
                if len(forcing) < n_steps * len(self.members):
                    raise ValueError("forcing table shorter than n_steps")
                fo = forcing
            else:
                if len(forcing) != n_steps:
                    raise ValueError("forcing must hold one list per step")
                fo = self.forcing_array(forcing)
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_ensemble_run(self.handle, fo, int(n_steps),
                                                S._stream(self.device)), "r2d3c_ensemble_run")

    def close(self):
        if self.handle:
            with torch.cuda.device(self.device):
                _lib.lib().r2d3c_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- the batched workload: 24 x 256^2 + 12 x 512^2 + 4 x 1024^2 --------------
MIX = ((256, 24), (512, 12), (1024, 4))


def mixed_ensemble_spec(mix=MIX, omega_seed=1234):
    """Per-run settings of the batched workload: run r (in size order) has
    seed r, Omega ~ U[0, 10] from numpy's default_rng(omega_seed),
    nu = nuv = 1e-3 * 256/n, hnu = hnuv = 0.1 (nn = nnv = 1, mm = mmv = 2),
    forcing shell kf = n/32 for psi and vz with injection rate 1.0, and the
    fixed dt = 1e-3 * 256/n (the CFL number k_c u dt stays ~0.25 for u ~ 3)."""
    rng = np.random.default_rng(omega_seed)
    out, r = [], 0
    for n, count in mix:
        for _ in range(count):
            nu = 1e-3 * 256.0 / n
            out.append(dict(n=n, seed=r, omega=float(rng.uniform(0.0, 10.0)), nu=nu,
                            hnu=0.1, kf=n / 32.0, f0=1.0, dt=1e-3 * 256.0 / n))
            r += 1
    return out


def build(spec, device, schedules="auto"):
    """Allocate the runs of `spec` (random-phase initial conditions seeded by
    each run's seed) and return their Ensemble."""
    from .ic import random_phase_ic
    members = []
    for r in spec:
        ps, vz = random_phase_ic(r["n"], r["seed"])
        p = S.Params(r["nu"], r["hnu"], 1, 2, r["nu"], r["hnu"], 1, 2, r["omega"])
        members.append(Member(torch.from_numpy(ps).to(device), torch.from_numpy(vz).to(device),
                              p, r["dt"]))
    return Ensemble(members, schedules)


def forcing_table(spec, n_steps):
    """rand_force draws of every run for n_steps steps (forcing seed = run seed)."""
    from .ic import forcing_schedule
    per_run = []
    for r in spec:
        sch = forcing_schedule(r["kf"], r["f0"], r["f0"], r["seed"], n_steps, S.rand_force_draw)
        per_run.append([(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sch])
    return [[per_run[m][st] for m in range(len(spec))] for st in range(n_steps)]


def points(spec):
    return sum(r["n"] * r["n"] for r in spec)
