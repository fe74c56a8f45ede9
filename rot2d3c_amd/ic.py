"""Initial conditions and forcing schedules (host side).

random_phase_ic: the seeded random-phase start used by the benchmark
workloads (no reference counterpart: the reference's own fresh start is
reference_ic below).  omega-hat and vz-hat get amplitude k^-1 and uniform
random phases on the shells kmin <= |k| <= kmax, psi-hat = omega-hat/k^2, and the
fields are scaled so that energy(ps,1) = <|grad psi|^2> and energy(vz,0) =
<vz^2> take the requested values (1.0 each is "energy 0.5" in the sense
1/2 <|u|^2> = 1/2 <vz^2> = 0.5).  Arrays use the reference spectral layout
(n/2+1, n), unnormalised (n^2 x coefficient).
"""
import math

import numpy as np


def _wavenumbers(n):
    kx = np.arange(n // 2 + 1, dtype=np.float64)
    kx[n // 2] = -n // 2  # ka(n/2+1) = -n/2 (src/Rot2D3C.para.f90:220-223)
    ky = np.fft.fftfreq(n, 1.0 / n)
    return kx, ky


def _energy(a, kin, n):
    """sum two*|a|^2*k^(2 kin)/n^4 (the reference's `energy`, kin >= 0)."""
    kx, ky = _wavenumbers(n)
    ka2 = kx[:, None] ** 2 + ky[None, :] ** 2
    two = np.full(n // 2 + 1, 2.0)
    two[0] = 1.0
    return float((two[:, None] * np.abs(a) ** 2 * ka2 ** kin).sum() / float(n) ** 4)


def random_phase_ic(n, seed, kmin=1.0, kmax=5.0, e_psi=1.0, e_vz=1.0):
    rng = np.random.default_rng(seed)
    kx, ky = _wavenumbers(n)
    ka2 = kx[:, None] ** 2 + ky[None, :] ** 2
    k = np.sqrt(ka2)
    shell = (k >= kmin) & (k <= kmax)
    shape = (n // 2 + 1, n)
    w = np.where(shell, np.exp(1j * rng.uniform(0.0, 2 * math.pi, shape)) / np.where(shell, k, 1.0), 0)
    v = np.where(shell, np.exp(1j * rng.uniform(0.0, 2 * math.pi, shape)) / np.where(shell, k, 1.0), 0)
    # kx = 0 column: Hermitian in ky so that the real fields are real
    for a in (w, v):
        a[0, n // 2 + 1:] = np.conj(a[0, 1:n // 2][::-1])
        a[0, 0] = 0.0
        a[0, n // 2] = 0.0
    ps = np.where(ka2 > 0, w / np.where(ka2 > 0, ka2, 1.0), 0).astype(np.complex128)
    vz = v.astype(np.complex128)
    ps *= math.sqrt(e_psi / _energy(ps, 1, n))
    vz *= math.sqrt(e_vz / _energy(vz, 0, n))
    return np.ascontiguousarray(ps), np.ascontiguousarray(vz)


def forcing_schedule(kup, fp0, fz0, seed, steps, draw):
    """Per-step rand_force draws of the iflow=3 loop for a run of `steps`
    steps: the pre-loop pair advances both seeds (src/Rot2D3C.para.f90:377-380),
    then each step draws psi (seed, kin=1) and vz (seed+1, kin=0) forcing
    (:413-415).  The draws do not depend on the flow, so a whole run can be
    scheduled up front.  `draw(kup, seed) -> (kx, ky, phase, seed)`.
    Returns a list of (psi_draw, vz_draw) tuples of (kx, ky, phase, f0, kupk)."""
    s0, s1 = seed, seed + 1
    _, _, _, s0 = draw(kup, s0)
    _, _, _, s1 = draw(kup, s1)
    out = []
    for _ in range(steps):
        kx, ky, ph, s0 = draw(kup, s0)
        a = (kx, ky, ph, fp0, float(kup))
        kx, ky, ph, s1 = draw(kup, s1)
        b = (kx, ky, ph, fz0, 1.0)
        out.append((a, b))
    return out
