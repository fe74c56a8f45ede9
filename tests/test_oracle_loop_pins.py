"""CPU pins of the oracle's main loop, fresh-start initial condition and iflow = 1
forcing against closed forms the mathematics fixes (tests/closed_forms.py):
each run starts from a state whose evolution is known exactly, so the output
schedule, the time written with each output, the restart counters and every
written value are checked without restating the loop."""
import math

import numpy as np
import pytest

import closed_forms as CF
from oracle import spectral as O

BASE = dict(CFL=0.4, step=9, cstep=1, sstep=3, tstep=4, fp0=0.0, fz0=0.0, u0=1.0, v0=0.7,
            kdn=2.0, kup=2.0, nu=1e-3, hnu=0.1, nn=1, mm=2, nuv=2e-2, hnuv=0.05, nnv=1, mmv=2,
            omega=0.0, seed=5, iflow=0, prm1=1.0, prm2=1.0, dt_corr=0.0, idir="ins", odir="outs")


def real(c):
    n = c.shape[1]
    return O.fftp2d_complex_to_real(c / float(n * n))


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - b) / (nb if nb > 0 else 1.0))


def check_rows(ev_bal, expected, tol=1e-10):
    assert len(ev_bal) == len(expected)
    for got, (e, z, h) in zip(ev_bal, expected):
        for key, want in (("energy_bal", e), ("enstrophy_bal", z), ("helicity_bal", h)):
            g = np.asarray(got[key])
            w = np.asarray(want)
            assert np.allclose(g, w, rtol=tol, atol=tol * max(1e-300, np.abs(w[1:]).max())), key


@pytest.mark.parametrize("n,kup,seed", [(32, 3.0, 5), (32, 2.0, 1234), (64, 5.0, 77)])
def test_initial_condition_pointwise(n, kup, seed):
    """src/Rot2D3C.para.f90:255-298: psi and vz of a fresh start equal the
    cosine sums evaluated point by point from the Park-Miller phases, scaled by
    the real-space Parseval norms, mean removed; the seed advances by
    4 (kup+1)^2 draws."""
    ps, vz, s = O.initial_condition(n, kup, 1.3, 0.7, seed)
    psi_e, vz_e, s_e = CF.ic_real(n, kup, 1.3, 0.7, seed, O.randu)
    assert s == s_e
    assert rel(real(ps), psi_e) < 1e-12
    assert rel(real(vz), vz_e) < 1e-12
    # energies after the mask: the mean of vz is removed, psi's has no energy
    assert math.isclose(O.energy(ps, 1), 1.3 ** 2, rel_tol=1e-12)
    assert O.energy(vz, 0) < 0.7 ** 2


@pytest.mark.parametrize("kup", [1.0, 2.0, 5.0])
def test_static_forcing_closed_form(kup):
    """src/Rot2D3C.para.f90:343-371: fp = (fp0/kup)(sin kup x + sin kup y) and
    fz = fz0 (cos kup x + sin kup y) in real space (the Parseval norms of the
    patterns are kup^2 and 1)."""
    n = 32
    fp, fz = O.static_forcing(n, kup, 0.8, 1.7)
    X, Y = CF.grid_xy(n)
    assert rel(real(fp), (0.8 / kup) * (np.sin(kup * X) + np.sin(kup * Y))) < 1e-13
    assert rel(real(fz), 1.7 * (np.cos(kup * X) + np.sin(kup * Y))) < 1e-13


def test_static_forcing_outside_the_disc_is_zero():
    fp, fz = O.static_forcing(32, 12.0, 0.8, 1.7)  # 12 > 32/3: only rounding leaks in
    assert np.abs(real(fp)).max() < 1e-13 and np.abs(real(fz)).max() < 1e-13


@pytest.mark.parametrize("step,cstep,sstep,tstep", [(9, 1, 3, 4), (7, 2, 5, 3)])
def test_reference_loop_fresh_start_closed_form(step, cstep, sstep, tstep):
    n = 32
    prm = dict(BASE, u0=0.0, omega=0.0, step=step, cstep=cstep, sstep=sstep, tstep=tstep)
    bal, fields, spec, vz_end, t_end = CF.fresh_decay_case(n, prm, 0.0, O.randu)
    ps, vz, time, ev = O.reference_loop(n, prm)
    assert not ps.any()
    assert rel(real(vz), vz_end) < 1e-12
    assert math.isclose(time, t_end, rel_tol=1e-13)
    check_rows(ev["balance"], bal)
    assert [e[0] for e in ev["field"]] == [f[0] for f in fields]
    for (idx, t, eps, evz), (_, te, vze) in zip(ev["field"], fields):
        assert math.isclose(t, te, rel_tol=1e-13, abs_tol=1e-15)
        assert not eps.any() and rel(real(evz), vze) < 1e-12
    assert [e[0] for e in ev["spectrum"]] == [s[0] for s in spec]
    for (idx, t, (ek, ezk), _), (_, te, (ek_e, ezk_e)) in zip(ev["spectrum"], spec):
        assert math.isclose(t, te, rel_tol=1e-13, abs_tol=1e-15)
        assert not ek.any() and np.allclose(ezk, ezk_e, rtol=1e-11, atol=1e-14 * ezk_e.max())


MODES = [(2, 3, 0.4 * np.exp(0.3j), 0.25 * np.exp(-1.1j)), (3, -1, 0.2 * np.exp(2.0j), 0.5)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("stat,step", [(3, 20), (2, 9)])
def test_reference_loop_restart_single_mode_closed_form(mode, stat, step):
    n = 32
    prm = dict(BASE, nu=0.0, hnu=0.0, nuv=0.0, hnuv=0.0, omega=2.5, CFL=0.3, step=step,
               cstep=2, sstep=3, tstep=4)
    time0 = 1.75
    bal, fields, spec, (psi_end, vz_end), t_end = CF.single_mode_case(n, prm, stat, time0, mode)
    psi0, vz0 = CF.SingleMode(n, *mode, prm["omega"]).fields()
    ps, vz, time, ev = O.reference_loop(n, prm, stat=stat, time=time0,
                                        ps=O.fftp2d_real_to_complex(psi0),
                                        vz=O.fftp2d_real_to_complex(vz0))
    assert rel(real(ps), psi_end) < 1e-12 and rel(real(vz), vz_end) < 1e-12
    assert math.isclose(time, t_end, rel_tol=1e-13)
    check_rows(ev["balance"], bal)
    assert [e[0] for e in ev["field"]] == [f[0] for f in fields]
    for (idx, t, eps, evz), (_, te, (pe, ve)) in zip(ev["field"], fields):
        assert math.isclose(t, te, rel_tol=1e-13)
        assert rel(real(eps), pe) < 1e-12 and rel(real(evz), ve) < 1e-12
    assert [e[0] for e in ev["spectrum"]] == [s[0] for s in spec]
    for (idx, t, (ek, ezk), (fl, fx)), (_, te, (ek_e, ezk_e)) in zip(ev["spectrum"], spec):
        assert math.isclose(t, te, rel_tol=1e-13)
        assert np.allclose(ek, ek_e, rtol=1e-11, atol=1e-13 * ek_e.max())
        assert np.allclose(ezk, ezk_e, rtol=1e-11, atol=1e-13 * ezk_e.max())
        for f in fx:  # one mode: both brackets vanish, so do the transfers
            assert np.abs(f).max() < 1e-12 * max(ek_e.max(), ezk_e.max())


def test_reference_loop_iflow1_closed_form():
    n = 32
    prm = dict(BASE, iflow=1, u0=0.0, v0=0.0, fz0=0.0, fp0=0.8, kup=3.0, kdn=1.0, omega=0.0,
               step=8, cstep=1, sstep=100, tstep=100, CFL=0.5)
    bal, psi_end, t_end = CF.iflow1_case(n, prm, 0.0)
    ps, vz, time, ev = O.reference_loop(n, prm)
    assert not vz.any()
    assert rel(real(ps), psi_end) < 1e-12
    assert math.isclose(time, t_end, rel_tol=1e-13)
    check_rows(ev["balance"], bal)
