"""Numpy restatement of the Rot2D3C hot path (CPU oracle, test infrastructure only).

Conventions (all follow the reference byte-for-byte in layout):
  * spectral arrays are complex128 of shape (n//2+1, n): index [i][j] is the
    Fortran element C(j,i) of `DOUBLE COMPLEX C(n, ista:iend)` with one rank,
    i.e. the slow index i is kx = 0..n/2 and the fast index j is ky in FFT order
    (src/Rot2D3C.para.f90:102-117, src/fftp2D.f90:153-156);
  * real arrays are float64 of shape (n, n): index [j][i] is R(i,j), x fastest;
  * spectral values are unnormalised (n^2 times the Fourier coefficient), the
    FFTW 2 convention (src/fftw_f77.i:7-11).

Pinning (checks that do not restate the formula; see tests/test_oracle_*.py):
  c2r / r2c .............. brute-force DFT matrices at n <= 16; delta and cosine
                            closed forms; numpy.fft round trip
  derivk2 / laplak2 ....... analytic derivatives of single Fourier modes
  poisson ................. analytic Jacobian of single-mode products; Galerkin
                            invariants <psi,J(psi,w)> = <w,J(psi,w)> = 0
  energy / inerprod ....... Parseval against real-space means of analytic fields
  CFL_condition ........... analytic max|grad psi| of a single mode
  randu ................... Park-Miller minimal-standard known answer
                            (state 1 -> 1043618065 after 10000 draws); float32
                            rounding checked against exact rationals
  const_inj ............... injection <ps, f>_kin equals fp0 exactly (the random
                            part is orthogonal to ps); band support; conjugate
                            symmetry of the kx = 0 column; the random factors
                            recovered from the output are, slab by slab, the
                            known-answer Park-Miller stream from seed = rank
                            (P = 1, 2, 4)
  rand_force .............. real-space transform is the analytic cosine of the
                            drawn wavevector; amplitude from the injection rate
  rk_stages ............... linear inertial-wave limit equals the 4th-order Taylor
                            polynomial of the exact propagator (5th-order local
                            error against exp(dt L)); pure dissipation equals
                            backward Euler; forcing enters as dt*f; energy (and,
                            at Omega=0, enstrophy) conserved to second order in
                            dt when nu = nu+ = 0 (the scheme is fourth order for
                            linear problems only); a mutation test plants
                            plausible slips and checks that these pins fail
  spectrum / transfers .... sum_k E(k) equals energy(); sum_k T(k) = 0
  cond_check .............. each column equals the Parseval real-space value;
                            the energy, enstrophy and helicity budgets close
                            (to 1e-3 with steady forcing, statistically with
                            random forcing) as the reference's BalanceTest
                            script checks them; max_min per slab by hand
  initial_condition ....... the cosine sums evaluated point by point from the
                            Park-Miller phases, normalised by real-space
                            Parseval means (tests/test_oracle_loop_pins.py)
  static_forcing .......... real-space closed form (fp0/kup)(sin kup x +
                            sin kup y), fz0 (cos kup x + sin kup y)
  reference_loop /
  run_random_forcing ...... closed-form runs: a fresh start that only decays
                            (u0 = 0, Omega = 0: constant dt, backward-Euler
                            factor per mode), a restart from one rotating
                            Fourier mode (each step the fourth-order Taylor
                            polynomial of exp(dt L), CFL dt from the exact
                            state) and an iflow = 1 start growing the forcing
                            pattern; they fix the output schedule, output
                            numbers, times written, restart counters and
                            iteration count (mutation tests plant slips in
                            each); the forcing lag of cond_check is pinned by
                            the random-forcing budget closure.  Parity
                            unpinned: the two pre-loop rand_force draws that
                            only advance the seeds (:376-380) -- any offset of
                            the forcing sequence is equally valid
                            statistically, and only a run of the reference
                            would fix it (the reference cannot be built here:
                            it needs FFTW 2.x, absent from the image)
"""
import math

import numpy as np

PI = 3.1415926535897932384  # src/pseudospec2D_mod.f90:35 (pi literal)


# --------------------------------------------------------------------------
# grid constants: src/Rot2D3C.para.f90:213-228, src/pseudospec2D_mod.f90:26-47
# --------------------------------------------------------------------------
class Grid:
    """Wavenumber tables of one grid size n (the module variables `ka`, `ka2`,
    `kmax`, `tiny`; src/Rot2D3C.para.f90:213-228)."""

    def __init__(self, n):
        if n < 4 or n & (n - 1):
            raise ValueError("n must be a power of two >= 4")
        self.n = n
        self.kmax = (float(n) / 3.0) ** 2
        self.tiny = 0.000001
        ka = np.empty(n)
        for i in range(n // 2):  # src/Rot2D3C.para.f90:220-223
            ka[i] = float(i)
            ka[i + n // 2] = float(i - n // 2)
        self.ka = ka
        self.kx = ka[: n // 2 + 1].copy()  # half-plane index uses ka(i), i=1..n/2+1
        self.ky = ka.copy()
        self.ka2 = self.kx[:, None] ** 2 + self.ky[None, :] ** 2  # ka2(j,i)
        self.dealias = self.ka2 <= self.kmax
        self.two = np.full(n // 2 + 1, 2.0)
        self.two[0] = 1.0  # `two = 1` for i == 1 (src/pseudospec2D_Rot2D3C.f90:371-372)


_GRIDS = {}


def grid(n):
    g = _GRIDS.get(n)
    if g is None:
        g = _GRIDS[n] = Grid(n)
    return g


def _n_of(a):
    return a.shape[1]


# --------------------------------------------------------------------------
# 2D FFTs: src/fftp2D.f90:133-290
# --------------------------------------------------------------------------
# FFT threads for timing runs only (the CPU baseline of the benchmark): with a
# thread count set, the 2D transforms go through scipy.fft's pocketfft with
# that many workers, equal to numpy.fft's within rounding.  Tests keep None.
_FFT_WORKERS = None


def set_fft_workers(workers):
    """Thread count of the 2D transforms (None: numpy.fft, one thread)."""
    global _FFT_WORKERS
    _FFT_WORKERS = workers


def fftp2d_complex_to_real(c):
    """Unnormalised inverse 2D transform, e^{+i}.  Complex backward transform
    along ky for every kx column, then real c2r along x, which ignores the
    imaginary parts at kx = 0 and kx = n/2 (src/fftp2D.f90:251,286-287).
    numpy's irfft2 over [ky][kx] applies exactly that order."""
    n = _n_of(c)
    if _FFT_WORKERS:
        import scipy.fft
        return scipy.fft.irfft2(c.T, s=(n, n), workers=_FFT_WORKERS) * float(n * n)
    return np.fft.irfft2(c.T, s=(n, n)) * float(n * n)


def fftp2d_real_to_complex(r):
    """Unnormalised forward 2D transform, e^{-i}: real r2c along x, then
    complex along y; output transposed to [kx][ky] (src/fftp2D.f90:171-208)."""
    if _FFT_WORKERS:
        import scipy.fft
        return np.ascontiguousarray(scipy.fft.rfft2(r, workers=_FFT_WORKERS).T)
    return np.ascontiguousarray(np.fft.rfft2(r).T)


# --------------------------------------------------------------------------
# spectral operators: src/pseudospec2D_Rot2D3C.f90:6-127
# --------------------------------------------------------------------------
def derivk2(a, dir):
    """dir=1: i*kx*a, dir=2: i*ky*a, else copy; zero where ka2 > kmax
    (src/pseudospec2D_Rot2D3C.f90:6-68)."""
    g = grid(_n_of(a))
    if dir == 1:
        b = 1j * g.kx[:, None] * a
    elif dir == 2:
        b = 1j * g.ky[None, :] * a
    else:
        b = a.copy()
    b[~g.dealias] = 0.0
    return b


def laplak2(a):
    """b = -ka2*a, no mask (src/pseudospec2D_Rot2D3C.f90:71-95)."""
    g = grid(_n_of(a))
    return -g.ka2 * a


def inv_laplak2(a):
    """b = -a/ka2 where ka2 > tiny, else 0 (src/pseudospec2D_Rot2D3C.f90:98-127)."""
    g = grid(_n_of(a))
    b = np.zeros_like(a)
    m = g.ka2 > g.tiny
    b[m] = -a[m] / g.ka2[m]
    return b


def poisson(a, b):
    """Poisson bracket {a,b} = r2c[(da/dx db/dy - da/dy db/dx)/n^4]
    (src/pseudospec2D_Rot2D3C.f90:169-226).  The output mask at :218-224 can
    never fire (ka2 >= kmax and ka2 <= tiny) and is therefore omitted."""
    n = _n_of(a)
    r1 = fftp2d_complex_to_real(derivk2(a, 1))
    r2 = fftp2d_complex_to_real(derivk2(b, 2))
    r3 = r1 * r2
    r1 = fftp2d_complex_to_real(derivk2(a, 2))
    r2 = fftp2d_complex_to_real(derivk2(b, 1))
    r3 = (r3 - r1 * r2) / float(n) ** 4
    return fftp2d_real_to_complex(r3)


# --------------------------------------------------------------------------
# global reductions: src/pseudospec2D_Rot2D3C.f90:336-451
# --------------------------------------------------------------------------
def _ka2pow(ka2, kin):
    # Fortran `ka2**kin` with an integer exponent; 0**0 == 1
    return ka2 ** kin if kin >= 0 else np.where(ka2 > 0, ka2, 1.0) ** kin


def energy(a, kin):
    """sum two*|a|^2*ka2^kin/n^4; kin < 0 skips ka2 < tiny
    (src/pseudospec2D_Rot2D3C.f90:336-396)."""
    g = grid(_n_of(a))
    w = g.two[:, None] * np.abs(a) ** 2 * _ka2pow(g.ka2, kin) / float(g.n) ** 4
    if kin < 0:
        w = np.where(g.ka2 >= g.tiny, w, 0.0)
    return float(w.sum())


def inerprod(a, b, kin):
    """sum two*ka2^kin*Re(b*conj(a))/n^4 over ka2 <= kmax (and ka2 >= tiny when
    kin < 0) (src/pseudospec2D_Rot2D3C.f90:400-451)."""
    g = grid(_n_of(a))
    m = g.dealias if kin >= 0 else g.dealias & (g.ka2 >= g.tiny)
    w = g.two[:, None] * _ka2pow(g.ka2, kin) * (b * np.conj(a)).real / float(g.n) ** 4
    return float(np.where(m, w, 0.0).sum())


# --------------------------------------------------------------------------
# CFL time step: src/pseudospec2D_Rot2D3C.f90:840-899
# --------------------------------------------------------------------------
def cfl_max_grad2(ps):
    """max over the grid of |grad psi|^2 (n^4-scaled), :863-874."""
    r1 = fftp2d_complex_to_real(derivk2(ps, 1))
    r2 = fftp2d_complex_to_real(derivk2(ps, 2))
    return float((r1 * r1 + r2 * r2).max())


def dt_from_max(cfl, n, maxg2, nu, nn, nuv, nnv, omega):
    """dt = cfl/(kcut*umax + nu*kcut^(2nn) + nuv*kcut^(2nnv) + 2*omega), :892-895."""
    kcut = float(n) / 3.0
    nrm = float(n) ** 2
    tmp = math.sqrt(maxg2) / nrm
    return cfl / (kcut * tmp + nu * kcut ** (2 * nn) + nuv * kcut ** (2 * nnv) + 2 * omega)


def CFL_condition(cfl, ps, vz, nu, nn, nuv, nnv, omega):
    """The vz maximum of :879-888 is computed by the reference but unused."""
    return dt_from_max(cfl, _n_of(ps), cfl_max_grad2(ps), nu, nn, nuv, nnv, omega)


# --------------------------------------------------------------------------
# random numbers and forcing: src/pseudospec2D_mod.f90:50-73,
# src/pseudospec2D_Rot2D3C.f90:1063-1126
# --------------------------------------------------------------------------
_IA, _IM, _IQ, _IR, _MASK = 16807, 2147483647, 127773, 2836, 123459876
_AM = np.float32(1.0) / np.float32(_IM)  # REAL am = 1./im  (== 2**-31 in float32)


def randu(idum):
    """Park-Miller minimal standard with Schrage's factorisation and an XOR
    mask; float32 result in [-1,1).  Returns (value, new idum)
    (src/pseudospec2D_mod.f90:52-71)."""
    idum ^= _MASK
    k = idum // _IQ
    idum = _IA * (idum - k * _IQ) - _IR * k
    if idum < 0:
        idum += _IM
    r = _AM * np.float32(idum)
    r = (r - np.float32(0.5)) * np.float32(2.0)
    idum ^= _MASK
    return np.float32(r), idum


def rand_force_draw(kup, seed):
    """The host part of rand_force: theta from randu(seed) (advances seed),
    phase from randu(seed+1) on a temporary (does not advance seed), and the
    rounded wavevector (src/pseudospec2D_Rot2D3C.f90:1086-1093).
    Returns (kx, ky, phase, new_seed)."""
    r, seed = randu(seed)
    theta = float(r) * PI
    r2, _ = randu(seed + 1)
    phase = float(r2) * PI
    kx = float(math.floor(kup * math.cos(theta) + 0.5))
    ky = float(math.floor(kup * math.sin(theta) + 0.5))
    return kx, ky, phase, seed


def rand_force(kdn, kup, fp0, dt, seed, kin, n):
    """Dense forcing array with one random wavevector of length ~kup
    (src/pseudospec2D_Rot2D3C.f90:1063-1126).  Returns (fp, new seed).
    Bug-compatible placement: the half-plane index matches +-ky and the full
    index matches kx; only the kx=0 column gets a conjugate partner."""
    g = grid(n)
    kx, ky, phase, seed = rand_force_draw(kup, seed)
    tmp = math.sqrt(2.0 * fp0) / (math.sqrt(dt) * kup ** kin)
    val = complex(tmp * math.cos(phase), tmp * math.sin(phase)) * float(n) ** 2
    fp = np.zeros((n // 2 + 1, n), dtype=np.complex128)
    for i in range(1, n // 2 + 1):
        if g.kx[i] == ky or g.kx[i] == -ky:
            for j in range(n):
                if g.ky[j] == kx:
                    fp[i, j] = val
    if g.kx[0] == ky:
        for j in range(1, n // 2):
            if g.ky[j] == kx:
                fp[0, j] = val
                fp[0, n - j] = val.conjugate()
    return fp, seed


def slab(n1, n2, nprocs, irank):
    """0-based half-open [lo, hi) of rank irank's share of the 1-based range
    [n1, n2] (the even split of `range`, src/fftp.f90:17-47)."""
    iwork1, iwork2 = divmod(n2 - n1 + 1, nprocs)
    ista = irank * iwork1 + n1 + min(irank, iwork2)
    iend = ista + iwork1 - 1 + (1 if iwork2 > irank else 0)
    return ista - 1, iend


def const_inj(ps, kdn, kup, fp0, kin, nprocs=1):
    """Constant-injection forcing (iflow = 2, src/pseudospec2D_Rot2D3C.f90:977-1060):
    f = ps/(|ps|^2+1) on the band kdn^2 <= ka2 <= kup^2, rescaled by
    fp0/inerprod(ps, f, kin), plus i*randu*|k|*ps.  The kx = 0 column is filled
    for j = 2..n/2 and mirrored by conjugation.  With nprocs ranks each rank
    owns the kx columns of its slab (src/Rot2D3C.para.f90:102) and re-seeds
    randu with its rank on every call (:1032), drawing in its own loop order.
    Returns fp."""
    n = _n_of(ps)
    g = grid(n)
    band = (g.ka2 <= kup * kup) & (g.ka2 >= kdn * kdn)
    fp = np.zeros_like(ps)
    fp[1:] = np.where(band[1:], ps[1:] / (np.abs(ps[1:]) ** 2 + 1.0), 0.0)
    for j in range(1, n // 2):
        if band[0, j]:
            fp[0, j] = ps[0, j] / (abs(ps[0, j]) ** 2 + 1.0)
            fp[0, n - j] = np.conj(fp[0, j])
    efp = inerprod(ps, fp, kin)
    out = np.zeros_like(ps)
    for rank in range(nprocs):
        seed1 = rank
        lo, hi = slab(1, n // 2 + 1, nprocs, rank)
        for i in range(lo, hi):
            js = range(n) if i != 0 else range(1, n // 2)
            for j in js:
                if band[i, j]:
                    r, seed1 = randu(seed1)
                    tmp3 = float(r) * math.sqrt(g.ka2[i, j])
                    out[i, j] = fp[i, j] * fp0 / efp + 1j * tmp3 * ps[i, j]
                    if i == 0:
                        out[0, n - j] = np.conj(out[0, j])
    return out


def rk_stages_const_inj(ps, vz, dt, p, kdn, kup, fp0, fz0, ord=4, nprocs=1):
    """rk_stages with iflow = 2: fp = const_inj(C1), fz = const_inj(C2) at the
    top of every stage (src/Rot2D3C.para.f90:565-569)."""
    g = grid(_n_of(ps))
    upd = (g.ka2 <= g.kmax) & (g.ka2 > g.tiny)
    ka2 = np.where(upd, g.ka2, 1.0)
    c1, c2 = ps.copy(), vz.copy()
    for o in range(ord, 0, -1):
        fp = const_inj(c1, kdn, kup, fp0, 1, nprocs)
        fz = const_inj(c2, kdn, kup, fz0, 0, nprocs)
        c4 = laplak2(c1)
        c5 = poisson(c1, c4)
        c6 = poisson(c1, c2)
        c3 = derivk2(c1, 2)
        c4 = derivk2(c2, 2)
        tmp1 = dt / float(o)
        tmp2 = 1.0 / ka2
        tmp3 = 1.0 + (p.nu * ka2 ** p.nn + p.hnu * tmp2 ** p.mm) * tmp1
        n1 = (ps + ((-c5 + 2 * p.omega * c4) * tmp2 + fp) * tmp1) / tmp3
        tmp3 = 1.0 + (p.nuv * ka2 ** p.nnv + p.hnuv * tmp2 ** p.mmv) * tmp1
        n2 = (vz + (c6 + 2 * p.omega * c3 + fz) * tmp1) / tmp3
        c1 = np.where(upd, n1, 0.0)
        c2 = np.where(upd, n2, 0.0)
    return derivk2(c1, 0), derivk2(c2, 0)


# --------------------------------------------------------------------------
# the RK stages: src/Rot2D3C.para.f90:545-611
# --------------------------------------------------------------------------
class Params:
    """Physical parameters read from parameter.inp (src/Rot2D3C.para.f90:136-162)."""

    def __init__(self, nu, hnu, nn, mm, nuv, hnuv, nnv, mmv, omega):
        self.nu, self.hnu, self.nn, self.mm = float(nu), float(hnu), int(nn), int(mm)
        self.nuv, self.hnuv, self.nnv, self.mmv = float(nuv), float(hnuv), int(nnv), int(mmv)
        self.omega = float(omega)


def rk_stages(ps, vz, fp, fz, dt, p, ord=4):
    """One time step: C1=ps, C2=vz; for o=ord..1 the semi-implicit update of
    :578-599; then ps=derivk2(C1,0), vz=derivk2(C2,0) (:610-611)."""
    g = grid(_n_of(ps))
    upd = (g.ka2 <= g.kmax) & (g.ka2 > g.tiny)
    ka2 = np.where(upd, g.ka2, 1.0)
    c1 = ps.copy()
    c2 = vz.copy()
    for o in range(ord, 0, -1):
        c4 = laplak2(c1)
        c5 = poisson(c1, c4)
        c6 = poisson(c1, c2)
        c3 = derivk2(c1, 2)
        c4 = derivk2(c2, 2)
        tmp1 = dt / float(o)
        tmp2 = 1.0 / ka2
        tmp3 = 1.0 + (p.nu * ka2 ** p.nn + p.hnu * tmp2 ** p.mm) * tmp1
        n1 = (ps + ((-c5 + 2 * p.omega * c4) * tmp2 + fp) * tmp1) / tmp3
        tmp3 = 1.0 + (p.nuv * ka2 ** p.nnv + p.hnuv * tmp2 ** p.mmv) * tmp1
        n2 = (vz + (c6 + 2 * p.omega * c3 + fz) * tmp1) / tmp3
        c1 = np.where(upd, n1, 0.0)
        c2 = np.where(upd, n2, 0.0)
    return derivk2(c1, 0), derivk2(c2, 0)


# --------------------------------------------------------------------------
# diagnostics: src/pseudospec2D_Rot2D3C.f90:455-837
# --------------------------------------------------------------------------
_F32_201 = float(np.float32(2.01))  # `(2.01)*kup*kup` is a default-REAL literal (:559)


def _slab_extrema(r, nprocs):
    """(max, min) of max_min.txt: each rank reduces its rows [jsta, jend]
    (src/Rot2D3C.para.f90:103) and both are combined with MPI_MAX
    (src/pseudospec2D_Rot2D3C.f90:578,586), so the "min" is the largest of the
    slab minima."""
    n = r.shape[0]
    sl = [slab(1, n, nprocs, k) for k in range(nprocs)]
    return (max(float(r[lo:hi].max()) for lo, hi in sl),
            max(float(r[lo:hi].min()) for lo, hi in sl))


def cond_check(ps, vz, fp, fz, time, p, kup, nprocs=1):
    """Rows written to energy_bal.txt, enstrophy_bal.txt, helicity_bal.txt and
    max_min.txt (src/pseudospec2D_Rot2D3C.f90:455-612); the sums do not depend
    on the rank count, the extrema do (nprocs slabs of real-space rows)."""
    g = grid(_n_of(ps))
    n4 = float(g.n) ** 4
    Ep1 = energy(ps, 1)
    injp1 = inerprod(ps, fp, 1)
    Dsp = energy(ps, p.nn + 1)
    Hds = energy(ps, 1 - p.mm)
    Epv = energy(vz, 0)
    injz = inerprod(vz, fz, 0)
    Dspv = energy(vz, p.nnv)
    Hdsv = energy(vz, -p.mmv)
    coup = inerprod(vz, derivk2(ps, 2), 0)
    Ep2 = energy(ps, 2)
    injp2 = inerprod(ps, fp, 2)
    Dspw = energy(ps, p.nn + 2)
    Hdsw = energy(ps, 2 - p.mm)
    Eph = inerprod(vz, ps, 1)
    injh1 = inerprod(vz, fp, 1)
    injh2 = inerprod(ps, fz, 1)
    Dsph1 = inerprod(vz, ps, p.nn + 1)
    Hdsh1 = inerprod(vz, ps, 1 - p.mm)
    Dsph2 = inerprod(ps, vz, p.nnv + 1)
    Hdsh2 = inerprod(ps, vz, 1 - p.mmv)
    band = (g.ka2 <= _F32_201 * kup * kup) & (g.ka2 >= kup * kup)
    two = g.two[:, None]
    Efp = float(np.where(band, two * g.ka2 * np.abs(ps) ** 2 / n4, 0.0).sum())
    Efpz = float(np.where(band, two * np.abs(vz) ** 2 / n4, 0.0).sum())
    maxw, minw = _slab_extrema(fftp2d_complex_to_real(laplak2(ps)), nprocs)
    maxv, minv = _slab_extrema(fftp2d_complex_to_real(derivk2(vz, 0)), nprocs)
    om2 = 2 * p.omega
    return {
        "energy_bal": [time, Ep1, Epv, injp1, injz, p.nu * Dsp, p.nuv * Dspv,
                       p.hnu * Hds, p.hnuv * Hdsv, om2 * coup, Efp, Efpz],
        "enstrophy_bal": [time, Ep2, injp2, p.nu * Dspw, p.hnu * Hdsw],
        "helicity_bal": [time, -Eph, -injh1, -injh2, -p.nu * Dsph1, -p.nuv * Dsph2,
                         -p.hnu * Hdsh1, -p.hnuv * Hdsh2],
        "max_min": [time, maxw, minw, maxv, minv],
    }


def _shells(g):
    return np.floor(np.sqrt(g.ka2) + 0.5).astype(np.int64)  # int(sqrt(ka2)+.5)


def _shell_sum(g, w):
    kmn = _shells(g)
    m = (kmn > 0) & (kmn <= g.n // 2 + 1)
    out = np.zeros(g.n // 2 + 1)
    np.add.at(out, kmn[m] - 1, w[m])
    return out


def spectrum(ps, vz):
    """Shell spectra Ek(k), Ezk(k), k = 1..n/2+1 (src/pseudospec2D_Rot2D3C.f90:615-712)."""
    g = grid(_n_of(ps))
    n4 = float(g.n) ** 4
    two = g.two[:, None]
    ek = _shell_sum(g, two * g.ka2 * np.abs(ps) ** 2 / n4)
    ezk = _shell_sum(g, two * np.abs(vz) ** 2 / n4)
    return ek, ezk


def transfers(ps, vz):
    """Shell transfers Fl0 (enstrophy), Fl1 (2D energy), Fl2 (vz energy) and
    their cumulative fluxes (src/pseudospec2D_Rot2D3C.f90:715-837)."""
    g = grid(_n_of(ps))
    n4 = float(g.n) ** 4
    two = g.two[:, None]
    c2 = poisson(ps, laplak2(ps))
    c1 = poisson(ps, vz)
    fl0 = _shell_sum(g, two * g.ka2 * (ps * np.conj(c2)).real / n4)
    fl1 = _shell_sum(g, two * (ps * np.conj(c2)).real / n4)
    fl2 = _shell_sum(g, -two * (vz * np.conj(c1)).real / n4)
    return (fl0, fl1, fl2), (np.cumsum(fl0), np.cumsum(fl1), np.cumsum(fl2))


# --------------------------------------------------------------------------
# initial condition of a fresh start (stat = 0): src/Rot2D3C.para.f90:245-298
# --------------------------------------------------------------------------
def initial_condition(n, kup, u0, v0, seed):
    """Random-phase cosine sums for psi and vz, normalised so that
    energy(ps,1) = u0**2 and energy(vz,0) = v0**2, then dealiased.
    Returns (ps, vz, seed after the draws)."""
    g = grid(n)
    kk = int(kup + 1)
    x = np.arange(n, dtype=np.float64)  # dble(i)-1
    cx1, cy1, cx2, cy2, w = [], [], [], [], []
    for ir in range(1, kk + 1):
        for jr in range(1, kk + 1):
            kr = math.sqrt(float(ir * ir + jr * jr))
            ph = []
            for _ in range(4):
                r, seed = randu(seed)
                ph.append(float(r) * 2.0 * PI)
            ax = 2.0 * (ir - 1) * PI * x / float(n)
            ay = 2.0 * (jr - 1) * PI * x / float(n)
            cx1.append(np.cos(ax + ph[0]))
            cy1.append(np.cos(ay + ph[1]) / kr)
            cx2.append(np.sin(ax + ph[2]))
            cy2.append(np.cos(ay + ph[3]))
    R1 = np.array(cy1).T @ np.array(cx1)  # [y][x]
    R2 = np.array(cy2).T @ np.array(cx2)
    ps = fftp2d_real_to_complex(R1)
    vz = fftp2d_real_to_complex(R2)
    ener = energy(ps, 1)
    enerv = energy(vz, 0)
    keep = (g.ka2 <= g.kmax) & (g.ka2 >= g.tiny)
    ps = np.where(keep, (u0 / math.sqrt(ener)) * ps, 0.0)
    vz = np.where(keep, (v0 / math.sqrt(enerv)) * vz, 0.0)
    return ps, vz, seed


# --------------------------------------------------------------------------
# static forcing of iflow = 1: src/Rot2D3C.para.f90:343-371
# --------------------------------------------------------------------------
def static_forcing(n, kup, fp0, fz0):
    """fp from sin(2 kup pi x/n) + sin(2 kup pi y/n) normalised to
    energy(fp,1) = fp0**2, fz from cos(2 kup pi x/n) + sin(2 kup pi y/n)
    normalised to energy(fz,0) = fz0**2, both then kept only on
    tiny <= ka2 <= kmax (x = dble(i)-1 along the fast index)."""
    g = grid(n)
    x = np.arange(n, dtype=np.float64)
    sx = np.sin(2 * kup * PI * x / float(n))
    cx = np.cos(2 * kup * PI * x / float(n))
    R1 = sx[None, :] + sx[:, None]   # [y][x]: sin(..x..) + sin(..y..)
    R2 = cx[None, :] + sx[:, None]
    fp = fftp2d_real_to_complex(R1)
    fz = fftp2d_real_to_complex(R2)
    tmp = fp0 / math.sqrt(energy(fp, 1))
    tmp1 = fz0 / math.sqrt(energy(fz, 0))
    keep = (g.ka2 <= g.kmax) & (g.ka2 >= g.tiny)
    return np.where(keep, tmp * fp, 0.0), np.where(keep, tmp1 * fz, 0.0)


# --------------------------------------------------------------------------
# the main loop for iflow = 3 (src/Rot2D3C.para.f90:376-415,532-611)
# --------------------------------------------------------------------------
def run_random_forcing(ps, vz, steps, cfl, p, kdn, kup, fp0, fz0, seed,
                       time=0.0, fixed_dt=None, on_check=None, cstep=None):
    """Advance `steps` iterations of the reference loop with iflow = 3.

    Order per iteration: CFL dt from the pre-step state, cond_check (forcing
    of the previous step), new forcing, time += dt, the RK stages.  Before the
    loop rand_force runs twice (to advance both seeds) with the CFL dt of the
    initial state (:376-380).  `fixed_dt` replaces the CFL dt (a mode the
    reference does not have).  Returns (ps, vz, time, dts, seed, seed1)."""
    n = _n_of(ps)
    seed1 = seed + 1
    dt = (fixed_dt if fixed_dt is not None
          else CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega))
    fp, seed = rand_force(kdn, kup, fp0, dt, seed, 1, n)
    fz, seed1 = rand_force(kdn, kup, fz0, dt, seed1, 0, n)
    dts = []
    timec = cstep
    for _ in range(steps):
        dt = (fixed_dt if fixed_dt is not None
              else CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega))
        if cstep is not None and timec == cstep:
            timec = 0
            if on_check is not None:
                on_check(cond_check(ps, vz, fp, fz, time, p, kup))
        fp, seed = rand_force(kdn, kup, fp0, dt, seed, 1, n)
        fz, seed1 = rand_force(kdn, kup, fz0, dt, seed1, 0, n)
        if cstep is not None:
            timec += 1
        time = time + dt
        dts.append(dt)
        ps, vz = rk_stages(ps, vz, fp, fz, dt, p)
    return ps, vz, time, dts, seed, seed1


# --------------------------------------------------------------------------
# the main loop with its output cadence (src/Rot2D3C.para.f90:245-250,
# 300-308, 343-381, 393-615)
# --------------------------------------------------------------------------
def reference_loop(n, prm, stat=0, time=0.0, ps=None, vz=None):
    """Run the reference loop for iflow = 0, 1 or 3 and return its events:
    balance rows (every cstep), spectra (every sstep) and field dumps (every
    tstep), each with the output index and time the reference writes.
    prm: dict with the parameter.inp names.  A fresh start (stat = 0) builds
    the initial condition itself; a restart takes ps, vz."""
    p = Params(prm["nu"], prm["hnu"], prm["nn"], prm["mm"], prm["nuv"], prm["hnuv"],
               prm["nnv"], prm["mmv"], prm["omega"])
    seed = prm["seed"]
    if stat == 0:
        ps, vz, seed = initial_condition(n, prm["kup"], prm["u0"], prm["v0"], seed)
        ini, timet, timec, times = 1, prm["tstep"], prm["cstep"], prm["sstep"]
        out_idx, spec_idx = 0, 0
    else:
        ini = (stat - 1) * prm["tstep"]
        timet = timec = times = 0
        out_idx = stat
        spec_idx = int(float(ini) / float(prm["sstep"]) + 1)
    ev = {"balance": [], "spectrum": [], "field": []}
    zero = np.zeros_like(ps)
    fp, fz = zero, zero
    seed1 = seed + 1
    cfl = prm["CFL"]
    if prm["iflow"] == 1:  # (:343-371)
        fp, fz = static_forcing(n, prm["kup"], prm["fp0"], prm["fz0"])
    if prm["iflow"] == 3:
        dt = CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
        fp, seed = rand_force(prm["kdn"], prm["kup"], prm["fp0"], dt, seed, 1, n)
        fz, seed1 = rand_force(prm["kdn"], prm["kup"], prm["fz0"], dt, seed1, 0, n)
    for _ in range(ini, prm["step"] + 1):
        dt = CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
        if timec == prm["cstep"]:
            timec = 0
            ev["balance"].append(cond_check(ps, vz, fp, fz, time, p, prm["kup"]))
        if prm["iflow"] == 3:
            fp, seed = rand_force(prm["kdn"], prm["kup"], prm["fp0"], dt, seed, 1, n)
            fz, seed1 = rand_force(prm["kdn"], prm["kup"], prm["fz0"], dt, seed1, 0, n)
        if times == prm["sstep"]:
            times = 0
            spec_idx += 1
            ev["spectrum"].append((spec_idx, time, spectrum(ps, vz), transfers(ps, vz)))
        if timet == prm["tstep"]:
            timet = 0
            out_idx += 1
            ev["field"].append((out_idx, time, ps.copy(), vz.copy()))
        timet += 1
        times += 1
        timec += 1
        time = time + dt
        ps, vz = rk_stages(ps, vz, fp, fz, dt, p)
    return ps, vz, time, ev
