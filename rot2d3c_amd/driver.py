"""The Rot2D3C main program on one GPU (src/Rot2D3C.para.f90:88-639).

Reads status.inp and parameter.inp from the working directory, builds or
reads the initial state, and runs the time loop in the reference's order:

  CFL dt from the pre-step state -> cond_check every cstep (forcing of the
  previous step) -> new rand_force draws (iflow = 3) -> spectrum and
  transfers every sstep -> field dumps every tstep -> time += dt -> the four
  stages (src/Rot2D3C.para.f90:393-615).

The CFL step, the forcing amplitude and the stages run inside one
r2d3c_step call per iteration; diagnostics and dumps use the device
operators of rot2d3c_amd.spectral and are written in the reference's file
formats (rot2d3c_amd.io), so the reference's analysis scripts read them.

The grid size n is a compile-time constant of the reference
(src/pseudospec2D_mod.f90:11); here it is the --n option.  iflow = 2
(const_inj) recomputes the forcing at every stage from the stage's own state.

--nprocs P reproduces what a P-rank run of the reference writes, where that
depends on the rank count: one slab file per rank in each field dump, the
const_inj random numbers (each rank restarts randu at seed = rank on its kx
slab) and the max_min.txt extrema (MPI_MAX of the per-slab values, also for
the minima).  Everything else is independent of P up to rounding.

Usage: python -m rot2d3c_amd.driver --n 256 [--workdir DIR] [--nprocs P]
"""
import argparse
import math
import os
import sys

import numpy as np
import torch

from . import io as rio
from . import spectral as S

PI = 3.1415926535897932384  # src/pseudospec2D_mod.f90:35


def reference_ic(n, kup, seed):
    """Real-space fields of the fresh start (stat = 0): random-phase cosine
    products on the square [0, kup]^2 of wavenumbers, psi weighted by 1/|k|
    (src/Rot2D3C.para.f90:255-279).  Phases come from randu(seed) in the
    order phase1..phase4 for ir outer, jr inner.  Returns (R1, R2, seed)."""
    kk = int(kup + 1)
    xs = np.arange(n, dtype=np.float64)
    cx1, cy1, cx2, cy2 = [], [], [], []
    for ir in range(1, kk + 1):
        for jr in range(1, kk + 1):
            kr = math.sqrt(float(ir * ir + jr * jr))
            ph = []
            for _ in range(4):
                r, seed = S.randu(seed)
                ph.append(r * 2.0 * PI)
            ax = 2.0 * (ir - 1) * PI * xs / float(n)
            ay = 2.0 * (jr - 1) * PI * xs / float(n)
            cx1.append(np.cos(ax + ph[0]))
            cy1.append(np.cos(ay + ph[1]) / kr)
            cx2.append(np.sin(ax + ph[2]))
            cy2.append(np.cos(ay + ph[3]))
    R1 = np.asarray(cy1).T @ np.asarray(cx1)  # [y][x]
    R2 = np.asarray(cy2).T @ np.asarray(cx2)
    return R1, R2, seed


def _masked_scale(a, scale, n, strict_dc=True):
    """a*scale where tiny <= ka2 <= kmax, zero elsewhere (src/Rot2D3C.para.f90:288-297)."""
    kx = torch.arange(n // 2 + 1, device=a.device, dtype=torch.float64)
    kx[n // 2] = -n / 2
    ky = torch.fft.fftfreq(n, 1.0 / n, device=a.device, dtype=torch.float64)
    ka2 = kx[:, None] ** 2 + ky[None, :] ** 2
    keep = (ka2 <= (n / 3.0) ** 2) & (ka2 >= 0.000001)
    return torch.where(keep, a * scale, torch.zeros_like(a))


class Run:
    def __init__(self, n, workdir=".", device="cuda:0", nprocs=1, log=None):
        self.n = int(n)
        self.wd = workdir
        self.dev = torch.device(device)
        self.nprocs = int(nprocs)
        if not 1 <= self.nprocs <= self.n // 2 + 1:
            raise ValueError("nprocs must be in 1..n/2+1")
        self.log = log or (lambda *a: None)
        self.stat, self.time = rio.read_status_inp(os.path.join(workdir, "status.inp"))
        self.prm = rio.read_parameter_inp(os.path.join(workdir, "parameter.inp"))
        q = self.prm
        self.params = S.Params(q["nu"], q["hnu"], q["nn"], q["mm"], q["nuv"], q["hnuv"],
                               q["nnv"], q["mmv"], q["omega"])
        if q["iflow"] not in (0, 1, 2, 3):
            raise ValueError("iflow = %d is not one of 1, 2, 3" % q["iflow"])
        self.odir = self._path(q["odir"])
        self.idir = self._path(q["idir"])

    def _path(self, d):
        return d if os.path.isabs(d) else os.path.join(self.wd, d)

    # --- initial state ---------------------------------------------------
    def initial_state(self):
        n, q = self.n, self.prm
        if self.stat == 0:
            R1, R2, self.seed = reference_ic(n, q["kup"], q["seed"])
            ps = S.fftp2d_real_to_complex(torch.from_numpy(R1).to(self.dev))
            vz = S.fftp2d_real_to_complex(torch.from_numpy(R2).to(self.dev))
            ener, enerv = S.energy(ps, 1), S.energy(vz, 0)
            ps = _masked_scale(ps, q["u0"] / math.sqrt(ener), n).contiguous()
            vz = _masked_scale(vz, q["v0"] / math.sqrt(enerv), n).contiguous()
        else:  # restart: real-space dumps of output `stat`, no masking (:322-329)
            self.seed = q["seed"]
            R1 = rio.read_field(self.idir, "ps", self.stat, n, self.nprocs)
            R2 = rio.read_field(self.idir, "vz", self.stat, n, self.nprocs)
            ps = S.fftp2d_real_to_complex(torch.from_numpy(R1).to(self.dev))
            vz = S.fftp2d_real_to_complex(torch.from_numpy(R2).to(self.dev))
        return ps, vz

    def static_forcing(self):
        """iflow = 1: fixed sin/cos forcing, normalised and dealiased
        (src/Rot2D3C.para.f90:343-371)."""
        n, q = self.n, self.prm
        xs = np.arange(n, dtype=np.float64)
        kup = q["kup"]
        sx = np.sin(2 * kup * PI * xs / float(n))
        cxs = np.cos(2 * kup * PI * xs / float(n))
        R1 = sx[None, :] + sx[:, None]
        R2 = cxs[None, :] + sx[:, None]
        fp = S.fftp2d_real_to_complex(torch.from_numpy(np.ascontiguousarray(R1)).to(self.dev))
        fz = S.fftp2d_real_to_complex(torch.from_numpy(np.ascontiguousarray(R2)).to(self.dev))
        eu, ev = S.energy(fp, 1), S.energy(fz, 0)
        fp = _masked_scale(fp, q["fp0"] / math.sqrt(eu), n).contiguous()
        fz = _masked_scale(fz, q["fz0"] / math.sqrt(ev), n).contiguous()
        return fp, fz

    # --- outputs ---------------------------------------------------------
    def write_fields(self, ps, vz, out):
        n = self.n
        inv = 1.0 / float(n) ** 2
        os.makedirs(self.odir, exist_ok=True)
        kx = torch.arange(n // 2 + 1, device=ps.device, dtype=torch.float64)
        kx[n // 2] = -n / 2
        ky = torch.fft.fftfreq(n, 1.0 / n, device=ps.device, dtype=torch.float64)
        ka2 = kx[:, None] ** 2 + ky[None, :] ** 2
        fields = {
            "ps": ps * inv,
            "vz": vz * inv,
            "ww": ps * ka2 * inv,
            "invlap_psi": S.inv_laplak2(ps) * inv,  # only valid if mm = 2 (:508-509)
        }
        for name, c in fields.items():
            r = S.fftp2d_complex_to_real(c.contiguous()).cpu().numpy()
            rio.write_field(self.odir, name, out, r, self.nprocs)
        with open(os.path.join(self.wd, "time_field.txt"), "a") as f:
            f.write("%03d%s\n" % (out, rio.fortran_f(self.time, 12, 6)))

    def write_spectra(self, ps, vz, ext4):
        os.makedirs(self.odir, exist_ok=True)
        ek, ezk = S.spectrum(ps, vz)
        rio.write_columns(os.path.join(self.odir, "spectrum.%s.txt" % ext4), (ek, ezk))
        fl, fx = S.transfers(ps, vz)
        rio.write_columns(os.path.join(self.odir, "transfer.%s.txt" % ext4), fl)
        rio.write_columns(os.path.join(self.odir, "fluxes.%s.txt" % ext4), fx)
        with open(os.path.join(self.wd, "time_spec.txt"), "a") as f:
            f.write("%s%s\n" % (ext4, rio.fortran_f(self.time, 12, 6)))

    def write_balance(self, ps, vz, fp, fz):
        rows = S.cond_check(ps, vz, fp, fz, self.time, self.params, self.prm["kup"], self.nprocs)
        for name in ("energy_bal", "enstrophy_bal", "helicity_bal", "max_min"):
            rio.append_rows(os.path.join(self.wd, name + ".txt"), [rows[name]])
        return rows

    # --- the loop --------------------------------------------------------
    def run(self):
        q, n = self.prm, self.n
        ps, vz = self.initial_state()
        step, tstep, sstep, cstep = q["step"], q["tstep"], q["sstep"], q["cstep"]
        if self.stat == 0:
            ini, timet, timec, times = 1, tstep, cstep, sstep
            out_idx, spec_idx = 0, 0
        else:
            ini = (self.stat - 1) * tstep
            timet = timec = times = 0
            out_idx = self.stat
            spec_idx = int(float(ini) / float(sstep) + 1)
        zero = torch.zeros_like(ps)
        dense = None
        ci = None
        fp_draw = fz_draw = None
        fpd = fzd = zero
        if q["iflow"] == 2:  # pre-loop const_inj pair (:372-375)
            ci = (q["kdn"], q["kup"], q["fp0"], q["fz0"], self.nprocs)
            fpd = S.const_inj(ps, q["kdn"], q["kup"], q["fp0"], 1, self.nprocs)
            fzd = S.const_inj(vz, q["kdn"], q["kup"], q["fz0"], 0, self.nprocs)
        if q["iflow"] == 1:
            fpd, fzd = self.static_forcing()
            dense = (fpd, fzd)
        seed = self.seed
        seed1 = seed + 1
        dt_dev = torch.zeros(1, dtype=torch.float64, device=self.dev)
        cfl = q["CFL"]
        if q["iflow"] == 3:  # pre-loop draws advance both seeds (:376-380)
            dt0 = S.CFL_condition(cfl, ps, vz, q["nu"], q["nn"], q["nuv"], q["nnv"], q["omega"])
            kx, ky, ph, seed = S.rand_force_draw(q["kup"], seed)
            fp_draw = S.ForceDraw(kx, ky, ph, q["fp0"], q["kup"])
            kx, ky, ph, seed1 = S.rand_force_draw(q["kup"], seed1)
            fz_draw = S.ForceDraw(kx, ky, ph, q["fz0"], 1.0)
            last_dt = dt0
        for t in range(ini, step + 1):
            if timec == cstep:
                timec = 0
                if q["iflow"] == 3:
                    fpd, fzd = S.forcing_fill(n, self.dev, fp_draw, fz_draw, last_dt)
                self.write_balance(ps, vz, fpd, fzd)
            if q["iflow"] == 3:
                kx, ky, ph, seed = S.rand_force_draw(q["kup"], seed)
                fp_draw = S.ForceDraw(kx, ky, ph, q["fp0"], q["kup"])
                kx, ky, ph, seed1 = S.rand_force_draw(q["kup"], seed1)
                fz_draw = S.ForceDraw(kx, ky, ph, q["fz0"], 1.0)
            if times == sstep:
                times = 0
                spec_idx += 1
                self.write_spectra(ps, vz, "%04d" % spec_idx)
            if timet == tstep:
                timet = 0
                out_idx += 1
                self.write_fields(ps, vz, out_idx)
            timet += 1
            times += 1
            timec += 1
            if ci is not None:
                # the forcing cond_check sees next is the one of the last stage
                # (o = 1), computed from that stage's input state
                S.rk4_step(ps, vz, self.params, cfl=cfl, dt_out=dt_dev, const_inj=ci)
                fpd, fzd = S.const_inj_last(ps)
            else:
                S.rk4_step(ps, vz, self.params, fp_draw, fz_draw, cfl=cfl, dt_out=dt_dev,
                           dense=dense)
            last_dt = float(dt_dev.item())
            if not math.isfinite(last_dt):
                raise FloatingPointError("non-finite dt at iteration %d: the run blew up" % t)
            self.time = self.time + last_dt  # time advances before the stages (:537)
            self.log("it %d  time %.6f  dt %.3e" % (t, self.time, last_dt))
        torch.cuda.synchronize(self.dev)
        return ps, vz


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, required=True, help="grid size (power of two)")
    ap.add_argument("--workdir", default=".")
    ap.add_argument("--device", default="cuda:0")
    ap.add_argument("--nprocs", "--nprocs-out", type=int, default=1,
                    help="MPI rank count to reproduce (slab files, const_inj "
                         "random numbers, max_min extrema)")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    r = Run(a.n, a.workdir, a.device, a.nprocs, log=print if a.verbose else None)
    r.run()
    return 0


if __name__ == "__main__":
    sys.exit(main())
