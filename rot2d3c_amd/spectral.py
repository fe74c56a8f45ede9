"""Reference-named operators of the Rot2D3C hot path on device tensors.

Each function mirrors one subroutine of the reference (same name, same
argument meaning, same conventions) and runs the HIP kernels of
librot2d3c.so through the C ABI; torch only holds the device memory and
supplies the current stream.

Arrays (one MPI rank of the reference, byte-identical layout):
  spectral: torch.complex128, shape (n//2+1, n), [kx][ky], unnormalised
  real:     torch.float64,    shape (n, n),      [y][x]
"""
import ctypes
import threading

import torch

from . import _lib
from ._lib import R2D3C_DT_CFL, R2D3C_DT_FIXED, check

_plans = {}
_plans_lock = threading.Lock()


class Plan:
    """Twiddles and step workspace for one grid size on one device
    (fftp2d_create_plan, src/fftp2D.f90:30-63)."""

    def __init__(self, n, device):
        self.n = int(n)
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib.lib().r2d3c_plan_create(ctypes.byref(h), self.n), "r2d3c_plan_create")
        self.handle = h

    @property
    def workspace_bytes(self):
        return int(_lib.lib().r2d3c_plan_workspace_bytes(self.handle))

    def close(self):
        if self.handle:
            with torch.cuda.device(self.device):
                torch.cuda.synchronize(self.device)
                _lib.lib().r2d3c_plan_destroy(self.handle)
            self.handle = None

    @staticmethod
    def get(n, device):
        device = torch.device(device)
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        key = (device.index, int(n))
        with _plans_lock:
            p = _plans.get(key)
            if p is None:
                p = _plans[key] = Plan(n, device)
            return p

    @staticmethod
    def release(n, device):
        """Destroy the cached plan of grid size n on device, if any (the
        n = 16384 plan holds ~41 GB of workspace)."""
        device = torch.device(device)
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        with _plans_lock:
            p = _plans.pop((device.index, int(n)), None)
        if p is not None:
            p.close()


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _spec_n(a):
    if a.dtype != torch.complex128 or not a.is_cuda or not a.is_contiguous() or a.dim() != 2:
        raise TypeError("spectral arrays must be contiguous complex128 CUDA tensors of shape (n/2+1, n)")
    n = a.shape[1]
    if a.shape[0] != n // 2 + 1:
        raise ValueError("spectral array shape %s is not (n/2+1, n)" % (tuple(a.shape),))
    return n


def _real_n(r):
    if r.dtype != torch.float64 or not r.is_cuda or not r.is_contiguous() or r.dim() != 2 \
            or r.shape[0] != r.shape[1]:
        raise TypeError("real arrays must be contiguous float64 CUDA tensors of shape (n, n)")
    return r.shape[0]


def _new_spec(n, like):
    return torch.empty((n // 2 + 1, n), dtype=torch.complex128, device=like.device)


def _new_real(n, like):
    return torch.empty((n, n), dtype=torch.float64, device=like.device)


# --- transforms: src/fftp2D.f90:133-290 -----------------------------------
def fftp2d_complex_to_real(c, out=None):
    """Unnormalised inverse 2D FFT (src/fftp2D.f90:214-290); input preserved."""
    n = _spec_n(c)
    out = _new_real(n, c) if out is None else out
    _real_n(out)
    p = Plan.get(n, c.device)
    check(_lib.lib().r2d3c_c2r(p.handle, _ptr(c), _ptr(out), _stream(c.device)), "r2d3c_c2r")
    return out


def fftp2d_real_to_complex(r, out=None):
    """Unnormalised forward 2D FFT (src/fftp2D.f90:133-211)."""
    n = _real_n(r)
    out = _new_spec(n, r) if out is None else out
    _spec_n(out)
    p = Plan.get(n, r.device)
    check(_lib.lib().r2d3c_r2c(p.handle, _ptr(r), _ptr(out), _stream(r.device)), "r2d3c_r2c")
    return out


# --- operators: src/pseudospec2D_Rot2D3C.f90:6-226 --------------------------
def derivk2(a, dir, out=None):
    """i*kx*a (dir 1), i*ky*a (dir 2) or a copy (dir 0), dealiased
    (src/pseudospec2D_Rot2D3C.f90:6-68)."""
    n = _spec_n(a)
    out = _new_spec(n, a) if out is None else out
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_derivk2(p.handle, _ptr(a), _ptr(out), int(dir), _stream(a.device)),
          "r2d3c_derivk2")
    return out


def laplak2(a, out=None):
    """-k^2 a (src/pseudospec2D_Rot2D3C.f90:71-95)."""
    n = _spec_n(a)
    out = _new_spec(n, a) if out is None else out
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_laplak2(p.handle, _ptr(a), _ptr(out), _stream(a.device)), "r2d3c_laplak2")
    return out


def inv_laplak2(a, out=None):
    """-a/k^2, zero at k = 0 (src/pseudospec2D_Rot2D3C.f90:98-127)."""
    n = _spec_n(a)
    out = _new_spec(n, a) if out is None else out
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_inv_laplak2(p.handle, _ptr(a), _ptr(out), _stream(a.device)),
          "r2d3c_inv_laplak2")
    return out


def poisson(a, b, out=None):
    """{a, b} (src/pseudospec2D_Rot2D3C.f90:169-226)."""
    n = _spec_n(a)
    if _spec_n(b) != n:
        raise ValueError("size mismatch")
    out = _new_spec(n, a) if out is None else out
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_poisson(p.handle, _ptr(a), _ptr(b), _ptr(out), _stream(a.device)),
          "r2d3c_poisson")
    return out


# --- reductions: src/pseudospec2D_Rot2D3C.f90:336-451 -----------------------
def energy_device(a, kin, out):
    n = _spec_n(a)
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_energy(p.handle, _ptr(a), int(kin), _ptr(out), _stream(a.device)),
          "r2d3c_energy")
    return out


def inerprod_device(a, b, kin, out):
    n = _spec_n(a)
    p = Plan.get(n, a.device)
    check(_lib.lib().r2d3c_inerprod(p.handle, _ptr(a), _ptr(b), int(kin), _ptr(out),
                                    _stream(a.device)), "r2d3c_inerprod")
    return out


def energy(a, kin):
    """Mean energy-like sum (src/pseudospec2D_Rot2D3C.f90:336-396); returns a float."""
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    return float(energy_device(a, kin, out).item())


def inerprod(a, b, kin):
    """Dealiased inner product (src/pseudospec2D_Rot2D3C.f90:400-451); returns a float."""
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    return float(inerprod_device(a, b, kin, out).item())


# --- parameters, CFL and the step ------------------------------------------
class Params:
    """Dissipation/rotation parameters (parameter.inp lines 12-20)."""

    def __init__(self, nu, hnu, nn, mm, nuv, hnuv, nnv, mmv, omega):
        self.nu, self.hnu, self.nn, self.mm = float(nu), float(hnu), int(nn), int(mm)
        self.nuv, self.hnuv, self.nnv, self.mmv = float(nuv), float(hnuv), int(nnv), int(mmv)
        self.omega = float(omega)

    def c(self):
        return _lib.Params(self.nu, self.hnu, self.nuv, self.hnuv, self.omega,
                           self.nn, self.mm, self.nnv, self.mmv)


def CFL_condition_device(cfl, ps, p, out):
    n = _spec_n(ps)
    pl = Plan.get(n, ps.device)
    cp = p.c()
    check(_lib.lib().r2d3c_cfl(pl.handle, ctypes.byref(cp), float(cfl), _ptr(ps), _ptr(out),
                               _stream(ps.device)), "r2d3c_cfl")
    return out


def CFL_condition(cfl, ps, vz, nu, nn, nuv, nnv, omega):
    """dt from the CFL condition (src/pseudospec2D_Rot2D3C.f90:840-899).  The
    vz maximum of the reference is unused, so vz is not read."""
    p = Params(nu, 0.0, nn, 0, nuv, 0.0, nnv, 0, omega)
    out = torch.empty(1, dtype=torch.float64, device=ps.device)
    return float(CFL_condition_device(cfl, ps, p, out).item())


class ForceDraw:
    """One rand_force draw: wavevector, phase, injection rate and kup**kin."""

    def __init__(self, kx, ky, phase, f0, kupk):
        self.kx, self.ky, self.phase = float(kx), float(ky), float(phase)
        self.f0, self.kupk = float(f0), float(kupk)

    def c(self):
        return _lib.ForceDraw(self.kx, self.ky, self.phase, self.f0, self.kupk, 1)


def _forcing_c(fp_draw, fz_draw, dense=None, const_inj=None):
    """fp_draw / fz_draw: ForceDraw or None; dense: optional (fp, fz) pair of
    static spectral arrays (either may be None); const_inj: optional
    (kdn, kup, fp0, fz0[, nprocs]) of iflow = 2."""
    f = _lib.Forcing()
    if const_inj is not None:
        f.ci_enabled = 1
        f.ci_kdn, f.ci_kup, f.ci_fp0, f.ci_fz0 = (float(x) for x in const_inj[:4])
        f.ci_nprocs = int(const_inj[4]) if len(const_inj) > 4 else 1
    if fp_draw is not None:
        f.psi = fp_draw.c()
    if fz_draw is not None:
        f.vz = fz_draw.c()
    if dense is not None:
        a, b = dense
        if a is not None:
            _spec_n(a)
            f.dense_psi = a.data_ptr()
        if b is not None:
            _spec_n(b)
            f.dense_vz = b.data_ptr()
    return f


def rk4_step(ps, vz, p, fp_draw=None, fz_draw=None, dt=None, cfl=None, dt_out=None,
             dense=None, const_inj=None):
    """One time step in place (src/Rot2D3C.para.f90:399 and :545-611).

    Exactly one of `dt` (fixed step) and `cfl` (CFL-adaptive step from the
    pre-step state, as the reference) must be given.  The dt used is written
    to `dt_out` (a 1-element float64 CUDA tensor) when given.  dense:
    optional (fp, fz) static forcing arrays added to the draws (iflow = 1).
    const_inj: (kdn, kup, fp0, fz0[, nprocs]) for iflow = 2 (forcing
    recomputed from the stage state at every stage, with the random numbers
    of an nprocs-rank run)."""
    n = _spec_n(ps)
    if _spec_n(vz) != n:
        raise ValueError("size mismatch")
    if (dt is None) == (cfl is None):
        raise ValueError("give exactly one of dt and cfl")
    pl = Plan.get(n, ps.device)
    cp = p.c()
    fo = _forcing_c(fp_draw, fz_draw, dense, const_inj)
    mode, val = (R2D3C_DT_FIXED, dt) if dt is not None else (R2D3C_DT_CFL, cfl)
    optr = _ptr(dt_out) if dt_out is not None else ctypes.c_void_p(0)
    check(_lib.lib().r2d3c_step(pl.handle, ctypes.byref(cp), _ptr(ps), _ptr(vz), ctypes.byref(fo),
                                mode, float(val), optr, _stream(ps.device)), "r2d3c_step")


def rk4_run(ps, vz, p, draws=None, nsteps=None, dt=None, cfl=None, dt_out=None):
    """`nsteps` time steps in place, bitwise equal to as many rk4_step calls.

    draws: sequence of (fp_draw, fz_draw) per step (None entries allowed), or
    None for no forcing.  dt_out: float64 CUDA tensor with at least nsteps
    elements, receiving each step's dt."""
    n = _spec_n(ps)
    if _spec_n(vz) != n:
        raise ValueError("size mismatch")
    if (dt is None) == (cfl is None):
        raise ValueError("give exactly one of dt and cfl")
    if nsteps is None:
        if draws is None:
            raise ValueError("give nsteps or draws")
        nsteps = len(draws)
    if draws is not None and len(draws) < nsteps:
        raise ValueError("fewer forcing draws than steps")
    if dt_out is not None and (dt_out.numel() < nsteps or dt_out.dtype != torch.float64):
        raise ValueError("dt_out needs nsteps float64 elements")
    pl = Plan.get(n, ps.device)
    cp = p.c()
    arr = None
    if draws is not None:
        arr = (_lib.Forcing * nsteps)()
        for k in range(nsteps):
            a, b = draws[k]
            arr[k] = _forcing_c(a, b)
    mode, val = (R2D3C_DT_FIXED, dt) if dt is not None else (R2D3C_DT_CFL, cfl)
    optr = _ptr(dt_out) if dt_out is not None else ctypes.c_void_p(0)
    check(_lib.lib().r2d3c_run(pl.handle, ctypes.byref(cp), _ptr(ps), _ptr(vz), arr, int(nsteps),
                               mode, float(val), optr, _stream(ps.device)), "r2d3c_run")


def forcing_fill(n, device, fp_draw, fz_draw, dt):
    """Dense fp, fz arrays of one draw, as rand_force fills them
    (src/pseudospec2D_Rot2D3C.f90:1100-1122)."""
    dev = torch.device(device)
    fp = torch.empty((n // 2 + 1, n), dtype=torch.complex128, device=dev)
    fz = torch.empty_like(fp)
    pl = Plan.get(n, dev)
    fo = _forcing_c(fp_draw, fz_draw)
    check(_lib.lib().r2d3c_forcing_fill(pl.handle, ctypes.byref(fo), float(dt), _ptr(fp), _ptr(fz),
                                        _stream(dev)), "r2d3c_forcing_fill")
    return fp, fz


def const_inj(ps, kdn, kup, fp0, kin, nprocs=1, out=None):
    """Constant-injection forcing of iflow = 2 (src/pseudospec2D_Rot2D3C.f90:977-1060)
    with the random numbers of an nprocs-rank run (randu restarted at seed =
    rank on each rank's kx slab)."""
    n = _spec_n(ps)
    out = _new_spec(n, ps) if out is None else out
    pl = Plan.get(n, ps.device)
    check(_lib.lib().r2d3c_const_inj(pl.handle, _ptr(ps), float(kdn), float(kup), float(fp0),
                                     int(kin), int(nprocs), _ptr(out), _stream(ps.device)),
          "r2d3c_const_inj")
    return out


def const_inj_last(ps):
    """(fp, fz) of the last stage of the most recent iflow = 2 step on ps's plan."""
    n = _spec_n(ps)
    pl = Plan.get(n, ps.device)
    fp, fz = _new_spec(n, ps), _new_spec(n, ps)
    check(_lib.lib().r2d3c_const_inj_last(pl.handle, _ptr(fp), _ptr(fz), _stream(ps.device)),
          "r2d3c_const_inj_last")
    return fp, fz


# --- diagnostics: src/pseudospec2D_Rot2D3C.f90:455-837 -----------------------
COND_NAMES = ("Ep1", "injp1", "Dsp", "Hds", "Epv", "injz", "Dspv", "Hdsv", "coup", "Ep2",
              "injp2", "Dspw", "Hdsw", "Eph", "injh1", "injh2", "Dsph1", "Hdsh1", "Dsph2",
              "Hdsh2", "Efp", "Efpz", "maxw", "minw", "maxv", "minv")


def cond_check_values(ps, vz, fp, fz, p, kup, nprocs=1):
    """The 26 device-computed quantities of cond_check as a dict of floats
    (src/pseudospec2D_Rot2D3C.f90:484-586); the extrema are those of an
    nprocs-rank run (MPI_MAX of per-slab values, also for the minima)."""
    n = _spec_n(ps)
    pl = Plan.get(n, ps.device)
    out = torch.empty(len(COND_NAMES), dtype=torch.float64, device=ps.device)
    cp = p.c()
    check(_lib.lib().r2d3c_cond_check(pl.handle, ctypes.byref(cp), float(kup), _ptr(ps), _ptr(vz),
                                      _ptr(fp), _ptr(fz), int(nprocs), _ptr(out),
                                      _stream(ps.device)),
          "r2d3c_cond_check")
    return dict(zip(COND_NAMES, out.tolist()))


def cond_check(ps, vz, fp, fz, time, p, kup, nprocs=1):
    """Rows of energy_bal.txt, enstrophy_bal.txt, helicity_bal.txt and
    max_min.txt, in the reference's column order (:592-608)."""
    v = cond_check_values(ps, vz, fp, fz, p, kup, nprocs)
    om2 = 2 * p.omega
    return {
        "energy_bal": [time, v["Ep1"], v["Epv"], v["injp1"], v["injz"], p.nu * v["Dsp"],
                       p.nuv * v["Dspv"], p.hnu * v["Hds"], p.hnuv * v["Hdsv"], om2 * v["coup"],
                       v["Efp"], v["Efpz"]],
        "enstrophy_bal": [time, v["Ep2"], v["injp2"], p.nu * v["Dspw"], p.hnu * v["Hdsw"]],
        "helicity_bal": [time, -v["Eph"], -v["injh1"], -v["injh2"], -p.nu * v["Dsph1"],
                         -p.nuv * v["Dsph2"], -p.hnu * v["Hdsh1"], -p.hnuv * v["Hdsh2"]],
        "max_min": [time, v["maxw"], v["minw"], v["maxv"], v["minv"]],
    }


def spectrum(ps, vz):
    """(Ek, Ezk) numpy arrays over shells k = 1..n/2+1 (:615-712)."""
    n = _spec_n(ps)
    pl = Plan.get(n, ps.device)
    out = torch.empty(2 * (n // 2 + 1), dtype=torch.float64, device=ps.device)
    check(_lib.lib().r2d3c_spectrum(pl.handle, _ptr(ps), _ptr(vz), _ptr(out), _stream(ps.device)),
          "r2d3c_spectrum")
    o = out.cpu().numpy()
    return o[: n // 2 + 1], o[n // 2 + 1:]


def transfers(ps, vz):
    """((Fl0, Fl1, Fl2), (fluxes)) over shells k = 1..n/2+1 (:715-837); the
    fluxes are the running sums from k = 1 (:816-831)."""
    import numpy as np
    n = _spec_n(ps)
    pl = Plan.get(n, ps.device)
    out = torch.empty(3 * (n // 2 + 1), dtype=torch.float64, device=ps.device)
    check(_lib.lib().r2d3c_transfers(pl.handle, _ptr(ps), _ptr(vz), _ptr(out), _stream(ps.device)),
          "r2d3c_transfers")
    o = out.cpu().numpy().reshape(3, n // 2 + 1)
    fl = (o[0], o[1], o[2])
    fx = []
    for f in fl:  # fx = Fl(i) + fx, in order
        acc, c = 0.0, np.empty_like(f)
        for i, x in enumerate(f):
            acc = x + acc
            c[i] = acc
        fx.append(c)
    return fl, tuple(fx)


# --- host random numbers: src/pseudospec2D_mod.f90:52-71 ---------------------
def randu(idum):
    """Returns (float32 value as float, new idum)."""
    s = ctypes.c_int(int(idum))
    v = _lib.lib().r2d3c_randu(ctypes.byref(s))
    return float(v), s.value


def rand_force_draw(kup, seed):
    """theta/phase/wavevector draw of rand_force; returns (kx, ky, phase, new seed)."""
    s = ctypes.c_int(int(seed))
    kx, ky, ph = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _lib.lib().r2d3c_rand_force_draw(float(kup), ctypes.byref(s), ctypes.byref(kx),
                                     ctypes.byref(ky), ctypes.byref(ph))
    return kx.value, ky.value, ph.value, s.value


def rand_force(kdn, kup, fp0, dt, seed, kin, n, device):
    """Dense forcing array of one draw and the advanced seed
    (src/pseudospec2D_Rot2D3C.f90:1063-1126)."""
    kx, ky, ph, seed = rand_force_draw(kup, seed)
    d = ForceDraw(kx, ky, ph, fp0, float(kup) ** kin)
    fp, _ = forcing_fill(n, device, d, None, dt)
    return fp, seed


def kmax_index(n):
    """Largest K with K^2 <= (n/3)^2: columns kx > K are zero after dealiasing."""
    kmax = (n / 3.0) ** 2
    k = 0
    while (k + 1) * (k + 1) <= kmax:
        k += 1
    return k


__all__ = [
    "Plan", "Params", "ForceDraw", "fftp2d_complex_to_real", "fftp2d_real_to_complex",
    "derivk2", "laplak2", "inv_laplak2", "poisson", "energy", "inerprod", "energy_device",
    "inerprod_device", "CFL_condition", "CFL_condition_device", "rk4_step", "rk4_run",
    "forcing_fill",
    "randu", "rand_force_draw", "rand_force", "kmax_index", "cond_check", "cond_check_values",
    "spectrum", "transfers", "COND_NAMES", "const_inj", "const_inj_last",
]
