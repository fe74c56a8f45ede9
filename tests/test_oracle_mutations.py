"""The oracle pins have teeth: plant plausible mistakes in a copy of the
oracle (on the CPU, in a temporary directory) and check that
tests/test_oracle_pins.py (operators, scheme, diagnostics; the last two
mistakes are caught only by its budget-closure pins) or
tests/test_oracle_loop_pins.py (main loop, initial condition, iflow = 1
forcing) fails for each one."""
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


LOOP_MUTATIONS = {
    "fresh-start output counter": (
        'ini, timet, timec, times = 1, prm["tstep"], prm["cstep"], prm["sstep"]',
        'ini, timet, timec, times = 1, 0, prm["cstep"], prm["sstep"]'),
    "restart first iteration": ('        ini = (stat - 1) * prm["tstep"]',
                                '        ini = stat * prm["tstep"]'),
    "restart spectrum index": ('spec_idx = int(float(ini) / float(prm["sstep"]) + 1)',
                               'spec_idx = int(float(ini) / float(prm["sstep"]))'),
    "IC phase pairing": ("cx1.append(np.cos(ax + ph[0]))", "cx1.append(np.cos(ax + ph[1]))"),
    "IC axes": ("R1 = np.array(cy1).T @ np.array(cx1)  # [y][x]",
                "R1 = np.array(cx1).T @ np.array(cy1)  # [y][x]"),
    "IC normalisation": ("    ener = energy(ps, 1)\n    enerv", "    ener = energy(ps, 0)\n    enerv"),
    "static forcing pattern": ("R2 = cx[None, :] + sx[:, None]", "R2 = sx[None, :] + cx[:, None]"),
    "CFL dt after the outputs": (
        '    for _ in range(ini, prm["step"] + 1):\n        dt = CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)',
        '    dt = CFL_condition(cfl, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)\n'
        '    for _ in range(ini, prm["step"] + 1):'),
}


def _run_mutated(tmp_path, before, after, test_files):
    src = open(os.path.join(ROOT, "oracle", "spectral.py")).read()
    assert before in src, before
    shutil.copytree(os.path.join(ROOT, "oracle"), tmp_path / "oracle")
    os.makedirs(tmp_path / "rot2d3c_amd")
    for f in ("__init__.py", "_lib.py", "balance.py"):  # the budget restatement
        shutil.copy(os.path.join(ROOT, "rot2d3c_amd", f), tmp_path / "rot2d3c_amd" / f)
    os.makedirs(tmp_path / "tests")
    for f in ("conftest.py", "closed_forms.py") + tuple(test_files):
        shutil.copy(os.path.join(ROOT, "tests", f), tmp_path / "tests" / f)
    (tmp_path / "oracle" / "spectral.py").write_text(src.replace(before, after))
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           *["tests/" + f for f in test_files]], cwd=tmp_path,
                          capture_output=True, text=True)


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(name, tmp_path):
    before, after = MUTATIONS[name]
    r = _run_mutated(tmp_path, before, after, ["test_oracle_pins.py"])
    assert r.returncode != 0, "mutation '%s' not caught:\n%s" % (name, r.stdout[-2000:])


@pytest.mark.parametrize("name", sorted(LOOP_MUTATIONS))
def test_planted_loop_mistake_is_caught(name, tmp_path):
    before, after = LOOP_MUTATIONS[name]
    r = _run_mutated(tmp_path, before, after, ["test_oracle_loop_pins.py"])
    assert r.returncode != 0, "mutation '%s' not caught:\n%s" % (name, r.stdout[-2000:])
