"""GPU: the driver's files against closed forms (tests/closed_forms.py).  Each
run starts from a state whose evolution the mathematics fixes: a fresh start
that only decays (u0 = 0, Omega = 0), a restart from one rotating Fourier mode
(inviscid, so each step is the fourth-order Taylor polynomial of the linear
rotation operator) and an iflow = 1 start that only grows the forcing pattern.
Checked: which outputs are written and their numbers, the time written with
each, the balance rows, the dumped fields and the spectra."""
import math
import os

import numpy as np
import pytest

import closed_forms as CF
from oracle import spectral as O

pytestmark = pytest.mark.gpu

BASE = dict(CFL=0.4, step=9, cstep=1, sstep=3, tstep=4, fp0=0.0, fz0=0.0, u0=1.0, v0=0.7,
            kdn=2.0, kup=2.0, nu=1e-3, hnu=0.1, nn=1, mm=2, nuv=2e-2, hnuv=0.05, nnv=1, mmv=2,
            omega=0.0, seed=5, iflow=0, prm1=1.0, prm2=1.0, dt_corr=0.0, idir="ins", odir="outs")


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - b) / (nb if nb > 0 else 1.0))


def _run(tmp_path, prm, stat, time0, n):
    from rot2d3c_amd import driver, io as rio
    rio.write_parameter_inp(str(tmp_path / "parameter.inp"), prm)
    rio.write_status_inp(str(tmp_path / "status.inp"), stat, time0)
    r = driver.Run(n, str(tmp_path), "cuda:0")
    ps, vz = r.run()
    return r, ps.cpu().numpy(), vz.cpu().numpy()


def _real(c):
    n = c.shape[1]
    return O.fftp2d_complex_to_real(c / float(n * n))


def _check_balance(tmp_path, bal, tol=1e-10):
    files = [np.atleast_2d(np.loadtxt(str(tmp_path / f)))
             for f in ("energy_bal.txt", "enstrophy_bal.txt", "helicity_bal.txt")]
    assert all(f.shape[0] == len(bal) for f in files), ([f.shape for f in files], len(bal))
    for k, rows in enumerate(bal):
        for f, want in zip(files, rows):
            w = np.asarray(want)
            assert np.allclose(f[k], w, rtol=tol, atol=tol * max(1e-300, np.abs(w[1:]).max())), k


def _check_fields(tmp_path, n, fields, time_tol=2e-6):
    from rot2d3c_amd import io as rio
    tf = open(str(tmp_path / "time_field.txt")).read().splitlines()
    assert [l[:3] for l in tf] == ["%03d" % f[0] for f in fields]
    for line, (idx, t, expect) in zip(tf, fields):
        assert abs(float(line[3:]) - t) < time_tol  # F12.6
        psi_e, vz_e = expect
        ps = rio.read_field(str(tmp_path / "outs"), "ps", idx, n)
        vz = rio.read_field(str(tmp_path / "outs"), "vz", idx, n)
        assert rel(ps, psi_e) < 1e-11 if np.abs(psi_e).max() > 0 else np.abs(ps).max() < 1e-14
        assert rel(vz, vz_e) < 1e-11 if np.abs(vz_e).max() > 0 else np.abs(vz).max() < 1e-14
        ww = rio.read_field(str(tmp_path / "outs"), "ww", idx, n)
        assert np.allclose(ww, -CF.lap(psi_e), rtol=1e-10, atol=1e-11 * max(1.0, np.abs(ww).max()))


def _check_spectra(tmp_path, spec):
    ts = open(str(tmp_path / "time_spec.txt")).read().splitlines()
    assert [l[:4] for l in ts] == ["%04d" % s[0] for s in spec]
    for line, (idx, t, (ek_e, ezk_e)) in zip(ts, spec):
        assert abs(float(line[4:]) - t) < 2e-6
        sp = np.loadtxt(str(tmp_path / "outs" / ("spectrum.%04d.txt" % idx)))
        scale = max(ek_e.max(), ezk_e.max())
        assert np.allclose(sp[:, 0], ek_e, rtol=1e-10, atol=1e-12 * scale)
        assert np.allclose(sp[:, 1], ezk_e, rtol=1e-10, atol=1e-12 * scale)


@pytest.mark.parametrize("step,cstep,sstep,tstep", [(9, 1, 3, 4), (7, 2, 5, 3)])
def test_driver_fresh_start_decay_closed_form(tmp_path, cuda, step, cstep, sstep, tstep):
    print("driver fresh decay n=32 step=%d" % step, flush=True)
    n = 32
    prm = dict(BASE, u0=0.0, omega=0.0, step=step, cstep=cstep, sstep=sstep, tstep=tstep)
    time0 = 0.5
    bal, fields, spec, vz_end, t_end = CF.fresh_decay_case(n, prm, time0, O.randu)
    r, ps, vz = _run(tmp_path, prm, 0, time0, n)
    assert np.abs(ps).max() == 0.0
    assert rel(_real(vz), vz_end) < 1e-12
    assert math.isclose(r.time, t_end, rel_tol=1e-13)
    _check_balance(tmp_path, bal)
    zero = np.zeros((n, n))
    _check_fields(tmp_path, n, [(i, t, (zero, v)) for i, t, v in fields])
    _check_spectra(tmp_path, spec)


@pytest.mark.parametrize("mode", [(2, 3, 0.4 * np.exp(0.3j), 0.25 * np.exp(-1.1j)),
                                  (3, -1, 0.2 * np.exp(2.0j), 0.5)])
def test_driver_restart_single_mode_closed_form(tmp_path, cuda, mode):
    print("driver restart single mode n=32 mode=%s" % (mode[:2],), flush=True)
    from rot2d3c_amd import io as rio
    n, stat = 32, 3
    prm = dict(BASE, nu=0.0, hnu=0.0, nuv=0.0, hnuv=0.0, omega=2.5, CFL=0.3, step=20,
               cstep=2, sstep=3, tstep=4)
    time0 = 1.75
    psi0, vz0 = CF.SingleMode(n, *mode, prm["omega"]).fields()
    os.makedirs(str(tmp_path / "ins"))
    rio.write_field(str(tmp_path / "ins"), "ps", stat, psi0)
    rio.write_field(str(tmp_path / "ins"), "vz", stat, vz0)
    bal, fields, spec, (psi_end, vz_end), t_end = CF.single_mode_case(n, prm, stat, time0, mode)
    r, ps, vz = _run(tmp_path, prm, stat, time0, n)
    assert rel(_real(ps), psi_end) < 1e-12 and rel(_real(vz), vz_end) < 1e-12
    assert math.isclose(r.time, t_end, rel_tol=1e-13)
    _check_balance(tmp_path, bal)
    _check_fields(tmp_path, n, fields)
    _check_spectra(tmp_path, spec)
    fx = np.loadtxt(str(tmp_path / "outs" / ("fluxes.%04d.txt" % spec[0][0])))
    assert np.abs(fx).max() < 1e-12 * max(spec[0][2][0].max(), spec[0][2][1].max())


def test_driver_iflow1_closed_form(tmp_path, cuda):
    print("driver iflow=1 closed form n=32", flush=True)
    n = 32
    prm = dict(BASE, iflow=1, u0=0.0, v0=0.0, fz0=0.0, fp0=0.8, kup=3.0, kdn=1.0, omega=0.0,
               step=8, cstep=1, sstep=100, tstep=100, CFL=0.5)
    bal, psi_end, t_end = CF.iflow1_case(n, prm, 0.0)
    r, ps, vz = _run(tmp_path, prm, 0, 0.0, n)
    assert np.abs(vz).max() == 0.0
    assert rel(_real(ps), psi_end) < 1e-12
    assert math.isclose(r.time, t_end, rel_tol=1e-13)
    _check_balance(tmp_path, bal)


def test_driver_iflow1_matches_oracle_loop(tmp_path, cuda):
    """A nonlinear iflow = 1 run (static forcing of psi and vz, rotation,
    random-phase start) through the driver against the oracle's loop."""
    print("driver iflow=1 vs oracle n=32", flush=True)
    n = 32
    prm = dict(BASE, iflow=1, fp0=0.9, fz0=0.6, kup=4.0, kdn=1.0, omega=1.5, step=9,
               cstep=1, sstep=3, tstep=4)
    r, ps, vz = _run(tmp_path, prm, 0, 0.0, n)
    ops, ovz, otime, ev = O.reference_loop(n, prm)
    assert rel(ps, ops) < 1e-11 and rel(vz, ovz) < 1e-11
    assert math.isclose(r.time, otime, rel_tol=1e-12)
    bal = np.loadtxt(str(tmp_path / "energy_bal.txt"))
    assert bal.shape[0] == len(ev["balance"])
    for row, b in zip(bal, ev["balance"]):
        assert np.allclose(row, b["energy_bal"], rtol=1e-10, atol=1e-13)
