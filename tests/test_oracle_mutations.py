"""The oracle pins have teeth: plant plausible mistakes in a copy of the
oracle (on the CPU, in a temporary directory) and check that
tests/test_oracle_pins.py fails for each one.  The last two are caught only
by the budget-closure pins."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

MUTATIONS = {
    "bracket sign": ("r3 = (r3 - r1 * r2) / float(n) ** 4", "r3 = (r1 * r2 - r3) / float(n) ** 4"),
    "derivative axis": ("b = 1j * g.kx[:, None] * a", "b = 1j * g.ky[None, :] * a"),
    "rotation sign": ("n2 = (vz + (c6 + 2 * p.omega * c3 + fz) * tmp1) / tmp3",
                      "n2 = (vz + (c6 - 2 * p.omega * c3 + fz) * tmp1) / tmp3"),
    "c2r normalisation": ("return np.fft.irfft2(c.T, s=(n, n)) * float(n * n)",
                          "return np.fft.irfft2(c.T, s=(n, n)) * float(n)"),
    "hypo-viscosity order": ("tmp3 = 1.0 + (p.nu * ka2 ** p.nn + p.hnu * tmp2 ** p.mm) * tmp1",
                             "tmp3 = 1.0 + (p.nu * ka2 ** p.nn + p.hnu * tmp2 ** p.nn) * tmp1"),
    "stage order": ("for o in range(ord, 0, -1):", "for o in range(1, ord + 1):"),
    "mask edge": ("self.dealias = self.ka2 <= self.kmax", "self.dealias = self.ka2 < self.kmax - 50"),
    # cond_check fed the forcing of the coming step instead of the last one
    "forcing lag": ("                on_check(cond_check(ps, vz, fp, fz, time, p, kup))",
                    "                on_check(cond_check(ps, vz, rand_force(kdn, kup, fp0, dt, seed, 1, n)[0],"
                    " rand_force(kdn, kup, fz0, dt, seed1, 0, n)[0], time, p, kup))"),
    "dissipation column": ("    Dspv = energy(vz, p.nnv)", "    Dspv = energy(vz, p.nnv + 1)"),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(name, tmp_path):
    before, after = MUTATIONS[name]
    src = open(os.path.join(ROOT, "oracle", "spectral.py")).read()
    assert before in src
    shutil.copytree(os.path.join(ROOT, "oracle"), tmp_path / "oracle")
    os.makedirs(tmp_path / "rot2d3c_amd")
    for f in ("__init__.py", "_lib.py", "balance.py"):  # the budget restatement
        shutil.copy(os.path.join(ROOT, "rot2d3c_amd", f), tmp_path / "rot2d3c_amd" / f)
    os.makedirs(tmp_path / "tests")
    for f in ("conftest.py", "test_oracle_pins.py"):
        shutil.copy(os.path.join(ROOT, "tests", f), tmp_path / "tests" / f)
    (tmp_path / "oracle" / "spectral.py").write_text(src.replace(before, after))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0, "mutation '%s' not caught:\n%s" % (name, r.stdout[-2000:])
