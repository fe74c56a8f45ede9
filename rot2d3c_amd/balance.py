"""Budget closure of the balance files, restated from the reference's
post-processing script (python_scripts/BalanceTest_Rot2D3C.py:10-164).

For each budget the script forms, from output `start` on,

    diff = dE/dt - (inj - diss),  dE = gradient(E)/2 (helicity: gradient(H)),
    dt = gradient(t),  inj halved for random (delta-correlated) forcing,

and prints mean(diff) with its standard error std(diff, ddof=1)/sqrt(len).
The budgets and their columns:
  energy_2d   energy_bal.txt  E = en,        inj = injp,        diss = nu Dsp + hnu Hds      (:16-35)
  energy_z    energy_bal.txt  E = enz,       inj = injz,        diss = nuv Dspv + hnuv Hdsv  (:46-66)
  energy      energy_bal.txt  E = en + enz,  inj = injp + injz, diss = all four              (:77-97)
  enstrophy   enstrophy_bal   E = Z,         inj = injp2,       diss = nu Dspw + hnu Hdsw    (:108-129)
  helicity    helicity_bal    E = H,         inj = -(injh1 + injh2), diss = all four         (:140-164)
The rotation coupling (column 10 of energy_bal.txt) only exchanges energy
between psi and vz, so with Omega != 0 only `energy` closes; enstrophy_bal
carries no rotation term at all.
"""
from __future__ import annotations

import os

import numpy as np

START = 10  # the script's start (:12)


def budget(t, e, inj, diss, start=START, random=True, half=True):
    """(mean, standard error) of dE/dt - inj + diss from output `start` on."""
    t = np.asarray(t, dtype=np.float64)[start:]
    e = np.asarray(e, dtype=np.float64)[start:]
    inj = np.asarray(inj, dtype=np.float64)[start:]
    diss = np.asarray(diss, dtype=np.float64)[start:]
    if random:
        inj = 0.5 * inj
    de = np.gradient(e) / (2.0 if half else 1.0)
    diff = de / np.gradient(t) - (inj - diss)
    return float(np.nanmean(diff)), float(np.nanstd(diff, ddof=1) / np.sqrt(len(diff)))


def budgets(energy_bal, enstrophy_bal, helicity_bal, start=START, random=True,
            script_helicity_sign=True):
    """All five budgets from the rows of the three files (2-D arrays, one row
    per output, columns as cond_check writes them).

    helicity_bal.txt holds -H, -injection and -dissipation
    (src/pseudospec2D_Rot2D3C.f90:602), which close as written; the script
    negates the injection once more (:146-147).  That only passes when the
    helicity injection averages to zero, as with independent random forcing
    of psi and vz.  script_helicity_sign=False leaves the columns as written."""
    t, en, enz, inj, injz, dis, disz, hdis, hdisz, _coup, _uf, _ufz = np.asarray(energy_bal).T
    tz, z, injw, disw, hdisw = np.asarray(enstrophy_bal).T
    th, h, inj1, inj2, d1, d2, hd1, hd2 = np.asarray(helicity_bal).T
    kw = dict(start=start, random=random)
    return {
        "energy_2d": budget(t, en, inj, dis + hdis, **kw),
        "energy_z": budget(t, enz, injz, disz + hdisz, **kw),
        "energy": budget(t, en + enz, inj + injz, dis + hdis + disz + hdisz, **kw),
        "enstrophy": budget(tz, z, injw, disw + hdisw, **kw),
        "helicity": budget(th, h, (-1.0 if script_helicity_sign else 1.0) * (inj1 + inj2),
                           d1 + d2 + hd1 + hd2, half=False, **kw),
    }


def budgets_from_files(path=".", start=START, random=True, script_helicity_sign=True):
    """The budgets of the balance files in directory `path`."""
    rd = lambda f: np.loadtxt(os.path.join(path, f), ndmin=2)
    return budgets(rd("energy_bal.txt"), rd("enstrophy_bal.txt"), rd("helicity_bal.txt"),
                   start=start, random=random, script_helicity_sign=script_helicity_sign)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="budget closure of the balance files")
    ap.add_argument("path", nargs="?", default=".")
    ap.add_argument("--start", type=int, default=START)
    ap.add_argument("--steady", action="store_true",
                    help="forcing constant in time (do not halve the injection)")
    a = ap.parse_args(argv)
    for name, (m, se) in budgets_from_files(a.path, a.start, not a.steady).items():
        print(f"{name:10s} mean(dE/dt - Inj + Diss) = {m:.6e} +/- {se:.6e}")


if __name__ == "__main__":
    main()
