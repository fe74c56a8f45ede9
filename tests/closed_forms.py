"""Closed forms that the mathematics fixes for the reference's main loop,
its fresh-start initial condition and its iflow = 1 forcing.  They pin the
oracle's loop (oracle.spectral.reference_loop, initial_condition,
static_forcing) and the GPU driver without restating either:

* fresh-start initial condition (src/Rot2D3C.para.f90:255-298): the cosine
  sums are evaluated point by point from the Park-Miller phases, and the
  normalisation is the real-space Parseval mean (energy(psi,1) = <|grad psi|^2>,
  energy(vz,0) = <vz^2>, both including the mean, which the mask then drops);
* u0 = 0, Omega = 0, no forcing: psi stays exactly zero, both brackets vanish,
  the CFL dt is the constant CFL/(nu kc^2nn + nuv kc^2nnv), and every vz mode
  is multiplied per step by 1/(1 + (nuv k^2nnv + hnuv k^-2mmv) dt): the
  scheme's stages all start from the step's initial state, so only the last
  stage (dt/1) survives for a linear dissipative term;
* one Fourier mode, nu = nu+ = 0: both brackets vanish (J(a cos, b cos) = 0 for
  a common wavevector), and for the linear rotation coupling
  d/dt (psi^, vz^) = (2 Omega i ky vz^/k^2, 2 Omega i ky psi^) of the README's
  equations the four-stage scheme is exactly the fourth-order Taylor
  polynomial of exp(dt L);
* iflow = 1 with u0 = v0 = fz0 = 0 and Omega = 0: psi stays a multiple
  A (sin(kup x) + sin(kup y)) of the forcing pattern (every mode on the one
  shell |k| = kup, so lap psi is a multiple of psi and both brackets vanish),
  and A advances per step as A' = (A + dt fp0/kup)/(1 + lambda dt), lambda =
  nu kup^2nn + hnu kup^-2mm; max|grad psi| = sqrt(2) kup A is reached at the
  grid point x = y = 0;
* the output schedule read from src/Rot2D3C.para.f90:245-250,302-308,393-537:
  a fresh start writes output j (j = 1, 2, ...) from the state after
  (j - 1) * tstep steps (likewise balance rows every cstep and spectra every
  sstep, starting at the initial state); a restart from output `stat` writes
  output stat + j after j * tstep steps, balance rows after j * cstep steps,
  spectra after j * sstep steps numbered from floor(ini/sstep + 1) + 1 with
  ini = (stat - 1) * tstep, and runs step - ini + 1 iterations; the time
  written is the time before the step (time += dt precedes the stages).

All sums below are taken over the full Fourier plane of numpy.fft.fft2
(Parseval: sum |F_k|^2 / n^4 = mean f^2), independently of the reference's
half-plane weights."""
import math

import numpy as np

F32_201 = float(np.float32(2.01))  # `(2.01)*kup*kup`, a default-REAL literal


def grid_xy(n):
    """Physical coordinates 2 pi (i-1)/n of the grid, as [y][x] arrays."""
    t = 2.0 * math.pi * np.arange(n) / n
    return np.meshgrid(t, t, indexing="xy")


def ic_phases(kup, seed, randu):
    """The phases of the fresh start, four randu draws per (ir, jr), ir outer."""
    out = []
    kk = int(kup + 1)
    for ir in range(1, kk + 1):
        for jr in range(1, kk + 1):
            ph = []
            for _ in range(4):
                r, seed = randu(seed)
                ph.append(float(r) * 2.0 * math.pi)
            out.append((ir - 1, jr - 1, math.sqrt(ir * ir + jr * jr), ph))
    return out, seed


def ic_real(n, kup, u0, v0, seed, randu):
    """Real-space psi and vz of the fresh start, evaluated point by point;
    normalised by the analytic Parseval means; mean removed (the mask keeps
    ka2 >= tiny).  Requires every mode inside the dealiasing disc."""
    assert 2 * kup * kup <= (n / 3.0) ** 2
    X, Y = grid_xy(n)
    R1 = np.zeros((n, n))
    R2 = np.zeros((n, n))
    G1x = np.zeros((n, n))
    G1y = np.zeros((n, n))
    phases, seed = ic_phases(kup, seed, randu)
    for a, b, kr, (p1, p2, p3, p4) in phases:
        cx, sx = np.cos(a * X + p1), np.sin(a * X + p1)
        cy, sy = np.cos(b * Y + p2), np.sin(b * Y + p2)
        R1 += cx * cy / kr
        G1x += -a * sx * cy / kr
        G1y += -b * cx * sy / kr
        R2 += np.sin(a * X + p3) * np.cos(b * Y + p4)
    ener = float(np.mean(G1x ** 2 + G1y ** 2))
    enerv = float(np.mean(R2 ** 2))
    psi = (u0 / math.sqrt(ener)) * (R1 - R1.mean())
    vz = (v0 / math.sqrt(enerv)) * (R2 - R2.mean())
    return psi, vz, seed


def full_k2(n):
    k = np.fft.fftfreq(n, 1.0 / n)
    return k[None, :] ** 2 + k[:, None] ** 2  # [ky][kx]


def kxy(n):
    k = np.fft.fftfreq(n, 1.0 / n)
    return k[None, :] + 0 * k[:, None], k[:, None] + 0 * k[None, :]  # KX, KY as [y][x]


def k2pow(k2, kin):
    out = np.zeros_like(k2)
    if kin >= 0:
        out = k2 ** kin
    else:
        m = k2 > 0
        out[m] = k2[m] ** kin
    return out


def energy(f, kin):
    """sum |F|^2 k^(2 kin) / n^4 over the full plane (k = 0 dropped for kin < 0)."""
    n = f.shape[0]
    F = np.fft.fft2(f)
    return float((np.abs(F) ** 2 * k2pow(full_k2(n), kin)).sum() / n ** 4)


def inner(a, b, kin):
    """sum Re(B conj A) k^(2 kin) / n^4 over the full plane."""
    n = a.shape[0]
    A, B = np.fft.fft2(a), np.fft.fft2(b)
    return float(((B * np.conj(A)).real * k2pow(full_k2(n), kin)).sum() / n ** 4)


def band_energy(f, kup, weight_k2):
    n = f.shape[0]
    k2 = full_k2(n)
    band = (k2 <= F32_201 * kup * kup) & (k2 >= kup * kup)
    F = np.fft.fft2(f)
    w = k2 if weight_k2 else 1.0
    return float((np.abs(F) ** 2 * w * band).sum() / n ** 4)


def deriv_y(f):
    n = f.shape[0]
    _, KY = kxy(n)
    return np.real(np.fft.ifft2(1j * KY * np.fft.fft2(f)))


def lap(f):
    n = f.shape[0]
    return np.real(np.fft.ifft2(-full_k2(n) * np.fft.fft2(f)))


def shell_sums(w_full):
    """Shell sums over kmn = int(sqrt(k2) + 0.5) in 1..n/2+1 of a full-plane weight."""
    n = w_full.shape[0]
    kmn = np.floor(np.sqrt(full_k2(n)) + 0.5).astype(int)
    out = np.zeros(n // 2 + 1)
    m = (kmn >= 1) & (kmn <= n // 2 + 1)
    np.add.at(out, kmn[m] - 1, w_full[m])
    return out


def spectra(psi, vz):
    n = psi.shape[0]
    P, V = np.fft.fft2(psi), np.fft.fft2(vz)
    return (shell_sums(full_k2(n) * np.abs(P) ** 2 / n ** 4),
            shell_sums(np.abs(V) ** 2 / n ** 4))


def balance_rows(psi, vz, time, prm, fp=None, fz=None):
    """energy_bal, enstrophy_bal and helicity_bal rows of a state with forcing
    fp, fz (real-space fields of the previous step's forcing; None = zero)."""
    nu, hnu, nn, mm = prm["nu"], prm["hnu"], prm["nn"], prm["mm"]
    nuv, hnuv, nnv, mmv = prm["nuv"], prm["hnuv"], prm["nnv"], prm["mmv"]
    om2, kup = 2 * prm["omega"], prm["kup"]
    fp = np.zeros_like(psi) if fp is None else fp
    fz = np.zeros_like(psi) if fz is None else fz
    e = [time, energy(psi, 1), energy(vz, 0), inner(psi, fp, 1), inner(vz, fz, 0),
         nu * energy(psi, nn + 1), nuv * energy(vz, nnv), hnu * energy(psi, 1 - mm),
         hnuv * energy(vz, -mmv), om2 * inner(vz, deriv_y(psi), 0),
         band_energy(psi, kup, True), band_energy(vz, kup, False)]
    z = [time, energy(psi, 2), inner(psi, fp, 2), nu * energy(psi, nn + 2),
         hnu * energy(psi, 2 - mm)]
    h = [time, -inner(vz, psi, 1), -inner(vz, fp, 1), -inner(psi, fz, 1),
         -nu * inner(vz, psi, nn + 1), -nuv * inner(psi, vz, nnv + 1),
         -hnu * inner(vz, psi, 1 - mm), -hnuv * inner(psi, vz, 1 - mmv)]
    return e, z, h


def dissipative_decay(f, steps, dt, nuv, hnuv, nnv, mmv):
    """Each Fourier mode of f times (1 + (nuv k^2nnv + hnuv k^-2mmv) dt)^-steps."""
    n = f.shape[0]
    k2 = full_k2(n)
    lam = nuv * k2pow(k2, nnv) + hnuv * k2pow(k2, -mmv)
    fac = (1.0 + lam * dt) ** (-steps)
    fac[k2 == 0] = 0.0
    return np.real(np.fft.ifft2(np.fft.fft2(f) * fac))


def taylor4(dt, L):
    """I + dt L + (dt L)^2/2 + (dt L)^3/6 + (dt L)^4/24."""
    A = dt * L
    out = np.eye(2, dtype=complex)
    term = np.eye(2, dtype=complex)
    for j in range(1, 5):
        term = term @ A / j
        out = out + term
    return out


class SingleMode:
    """psi = Re(a e^{i k.x}), vz = Re(b e^{i k.x}) on the n x n grid."""

    def __init__(self, n, kx, ky, a, b, omega):
        self.n, self.kx, self.ky = n, kx, ky
        self.a, self.b = complex(a), complex(b)
        k2 = float(kx * kx + ky * ky)
        w = 2.0 * omega * 1j * ky
        self.L = np.array([[0.0, w / k2], [w, 0.0]], dtype=complex)
        X, Y = grid_xy(n)
        self.phase = np.exp(1j * (kx * X + ky * Y))

    def fields(self):
        return np.real(self.a * self.phase), np.real(self.b * self.phase)

    def grad_max(self):
        """max over the grid of |grad psi| (the CFL maximum)."""
        gx = np.real(1j * self.kx * self.a * self.phase)
        gy = np.real(1j * self.ky * self.a * self.phase)
        return float(np.sqrt((gx ** 2 + gy ** 2).max()))

    def advance(self, dt):
        self.a, self.b = taylor4(dt, self.L) @ np.array([self.a, self.b])


def cfl_dt(cfl, n, umax, prm):
    """dt = CFL/(kc umax + nu kc^(2nn) + nuv kc^(2nnv) + 2 Omega), kc = n/3
    (src/pseudospec2D_Rot2D3C.f90:895)."""
    kc = n / 3.0
    return cfl / (kc * umax + prm["nu"] * kc ** (2 * prm["nn"]) +
                  prm["nuv"] * kc ** (2 * prm["nnv"]) + 2 * prm["omega"])


def fresh_schedule(step, every):
    """Steps after which a fresh start writes (state before iteration t = 1..step)."""
    return [m for m in range(0, step) if m % every == 0]


def restart_schedule(step, stat, tstep, every):
    """Steps after which a restart writes: multiples of `every` up to the last
    of the step - ini + 1 iterations, ini = (stat - 1) tstep."""
    ini = (stat - 1) * tstep
    count = step - ini + 1
    return [m for m in range(1, count) if m % every == 0]


# ---- expected events of the three closed-form runs ----------------------
def fresh_decay_case(n, prm, time0, randu):
    """Expected events of a fresh start with u0 = 0, Omega = 0, no forcing."""
    _, vz0, _ = ic_real(n, prm["kup"], 0.0, prm["v0"], prm["seed"], randu)
    kc = n / 3.0
    dt = prm["CFL"] / (prm["nu"] * kc ** (2 * prm["nn"]) + prm["nuv"] * kc ** (2 * prm["nnv"]))
    zero = np.zeros((n, n))

    def vz_after(m):
        return dissipative_decay(vz0, m, dt, prm["nuv"], prm["hnuv"], prm["nnv"], prm["mmv"])

    bal = [balance_rows(zero, vz_after(m), time0 + m * dt, prm)
           for m in fresh_schedule(prm["step"], prm["cstep"])]
    fields = [(j + 1, time0 + m * dt, vz_after(m))
              for j, m in enumerate(fresh_schedule(prm["step"], prm["tstep"]))]
    spec = [(j + 1, time0 + m * dt, spectra(zero, vz_after(m)))
            for j, m in enumerate(fresh_schedule(prm["step"], prm["sstep"]))]
    return bal, fields, spec, vz_after(prm["step"]), time0 + prm["step"] * dt


def single_mode_case(n, prm, stat, time0, mode):
    """Expected events of a restart from one Fourier mode (nu = nu+ = 0)."""
    kx, ky, a, b = mode
    sm = SingleMode(n, kx, ky, a, b, prm["omega"])
    ini = (stat - 1) * prm["tstep"]
    count = prm["step"] - ini + 1
    states, times = [], []
    t = time0
    for m in range(count + 1):
        states.append(sm.fields())
        times.append(t)
        if m < count:
            dt = cfl_dt(prm["CFL"], n, sm.grad_max(), prm)
            sm.advance(dt)
            t += dt
    bal = [balance_rows(*states[m], times[m], prm)
           for m in restart_schedule(prm["step"], stat, prm["tstep"], prm["cstep"])]
    fields = [(stat + j + 1, times[m], states[m])
              for j, m in enumerate(restart_schedule(prm["step"], stat, prm["tstep"], prm["tstep"]))]
    first_spec = int(float(ini) / float(prm["sstep"]) + 1)
    spec = [(first_spec + j + 1, times[m], spectra(*states[m]))
            for j, m in enumerate(restart_schedule(prm["step"], stat, prm["tstep"], prm["sstep"]))]
    return bal, fields, spec, states[count], times[count]


def iflow1_case(n, prm, time0):
    """Expected events of a fresh iflow = 1 start with u0 = v0 = fz0 = 0, Omega = 0."""
    kup, fp0 = prm["kup"], prm["fp0"]
    lam = prm["nu"] * kup ** (2 * prm["nn"]) + prm["hnu"] * kup ** (-2 * prm["mm"])
    X, Y = grid_xy(n)
    pat = np.sin(kup * X) + np.sin(kup * Y)
    fp = (fp0 / kup) * pat
    zero = np.zeros((n, n))
    amps, times = [0.0], [time0]
    for m in range(prm["step"]):
        A = amps[-1]
        dt = cfl_dt(prm["CFL"], n, math.sqrt(2.0) * kup * abs(A), prm)
        amps.append((A + dt * fp0 / kup) / (1.0 + lam * dt))
        times.append(times[-1] + dt)
    bal = [balance_rows(amps[m] * pat, zero, times[m], prm, fp=fp)
           for m in fresh_schedule(prm["step"], prm["cstep"])]
    return bal, amps[-1] * pat, times[-1]
