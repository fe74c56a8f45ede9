"""GPU: the driver (rot2d3c_amd.driver) runs the reference's main loop and
writes the reference's files; compare them with the oracle's loop."""
import math
import os

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu

PRM = dict(CFL=0.4, step=9, cstep=1, sstep=3, tstep=4, fp0=1.0, fz0=0.5, u0=1.0, v0=0.7,
           kdn=4.0, kup=4.0, nu=1e-3, hnu=0.1, nn=1, mm=2, nuv=2e-3, hnuv=0.05, nnv=1, mmv=2,
           omega=1.5, seed=5, iflow=3, prm1=1.0, prm2=1.0, dt_corr=0.0, idir="ins", odir="outs")


def _setup(tmp_path, stat=0, time=0.0):
    from rot2d3c_amd import io as rio
    rio.write_parameter_inp(str(tmp_path / "parameter.inp"), PRM)
    rio.write_status_inp(str(tmp_path / "status.inp"), stat, time)


def test_driver_fresh_start_matches_oracle_loop(tmp_path, cuda):
    print("driver fresh start n=32", flush=True)
    from rot2d3c_amd import driver, io as rio
    n = 32
    _setup(tmp_path)
    r = driver.Run(n, str(tmp_path), "cuda:0")
    gps, gvz = r.run()
    ops, ovz, otime, ev = O.reference_loop(n, PRM)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(gps.cpu().numpy(), ops) < 1e-11 and rel(gvz.cpu().numpy(), ovz) < 1e-11
    assert math.isclose(r.time, otime, rel_tol=1e-12)
    bal = np.loadtxt(str(tmp_path / "energy_bal.txt"))
    assert bal.shape[0] == len(ev["balance"]) == PRM["step"]
    for row, b in zip(bal, ev["balance"]):
        assert np.allclose(row, b["energy_bal"], rtol=1e-10, atol=1e-13)
    ens = np.loadtxt(str(tmp_path / "enstrophy_bal.txt"))
    hel = np.loadtxt(str(tmp_path / "helicity_bal.txt"))
    mm = np.loadtxt(str(tmp_path / "max_min.txt"))
    for k, b in enumerate(ev["balance"]):
        assert np.allclose(ens[k], b["enstrophy_bal"], rtol=1e-10, atol=1e-13)
        assert np.allclose(hel[k], b["helicity_bal"], rtol=1e-10, atol=1e-13)
        assert np.allclose(mm[k], b["max_min"], rtol=1e-10, atol=1e-10)
    for idx, t, (ek, ezk), ((f0, f1, f2), (x0, x1, x2)) in ev["spectrum"]:
        sp = np.loadtxt(str(tmp_path / "outs" / ("spectrum.%04d.txt" % idx)))
        assert np.allclose(sp[:, 0], ek, rtol=1e-10, atol=1e-14 * ek.max())
        fl = np.loadtxt(str(tmp_path / "outs" / ("fluxes.%04d.txt" % idx)))
        assert np.allclose(fl[:, 1], x1, rtol=1e-9, atol=1e-11 * np.abs(x1).max())
    ts = open(str(tmp_path / "time_spec.txt")).read().splitlines()
    assert [l[:4] for l in ts] == ["%04d" % e[0] for e in ev["spectrum"]]
    tf = open(str(tmp_path / "time_field.txt")).read().splitlines()
    assert [l[:3] for l in tf] == ["%03d" % e[0] for e in ev["field"]]
    for idx, t, ps, vz in ev["field"]:
        assert abs(float(tf[idx - 1][3:]) - t) < 1e-6
        R = rio.read_field(str(tmp_path / "outs"), "ps", idx, n)
        expect = O.fftp2d_complex_to_real(ps / float(n * n))
        assert np.allclose(R, expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())


def test_driver_restart_reads_its_own_dumps(tmp_path, cuda):
    """stat > 0: the fields of output `stat` are read back from idir and the
    counters restart as in src/Rot2D3C.para.f90:300-329."""
    print("driver restart n=32", flush=True)
    from rot2d3c_amd import driver, io as rio
    n = 32
    _setup(tmp_path)
    driver.Run(n, str(tmp_path), "cuda:0").run()
    os.rename(str(tmp_path / "outs"), str(tmp_path / "ins"))
    tf = open(str(tmp_path / "time_field.txt")).read().splitlines()
    t2 = float(tf[1][3:])
    rio.write_status_inp(str(tmp_path / "status.inp"), 2, t2)
    for f in ("energy_bal.txt", "time_field.txt", "time_spec.txt"):
        os.remove(str(tmp_path / f))
    r = driver.Run(n, str(tmp_path), "cuda:0")
    gps, gvz = r.run()
    R1 = rio.read_field(str(tmp_path / "ins"), "ps", 2, n)
    R2 = rio.read_field(str(tmp_path / "ins"), "vz", 2, n)
    ps0, vz0 = O.fftp2d_real_to_complex(R1), O.fftp2d_real_to_complex(R2)
    ops, ovz, otime, ev = O.reference_loop(n, PRM, stat=2, time=t2, ps=ps0, vz=vz0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(gps.cpu().numpy(), ops) < 1e-11
    tf2 = open(str(tmp_path / "time_field.txt")).read().splitlines()
    assert [l[:3] for l in tf2] == ["%03d" % e[0] for e in ev["field"]]


@pytest.mark.parametrize("nprocs", [1, 2])
def test_driver_iflow2_matches_oracle_loop(tmp_path, cuda, nprocs):
    """iflow = 2 through the driver: the stage-by-stage step and the balance
    rows that see each step's last-stage forcing.  nprocs = 2 reproduces a
    2-rank run: const_inj random numbers per slab, max_min per slab."""
    print("driver iflow=2 n=32 nprocs=%d" % nprocs, flush=True)
    from rot2d3c_amd import driver, io as rio
    n = 32
    # with 2 ranks the band reaches into rank 1's kx slab (kx >= 9)
    prm = dict(PRM, iflow=2, kdn=2.0, kup=5.0 if nprocs == 1 else 10.0, step=6, sstep=100,
               tstep=100)
    rio.write_parameter_inp(str(tmp_path / "parameter.inp"), prm)
    rio.write_status_inp(str(tmp_path / "status.inp"), 0, 0.0)
    r = driver.Run(n, str(tmp_path), "cuda:0", nprocs=nprocs)
    gps, gvz = r.run()
    # oracle loop with per-stage const_inj
    p = O.Params(prm["nu"], prm["hnu"], prm["nn"], prm["mm"], prm["nuv"], prm["hnuv"],
                 prm["nnv"], prm["mmv"], prm["omega"])
    ps, vz, _ = O.initial_condition(n, prm["kup"], prm["u0"], prm["v0"], prm["seed"])
    fp = O.const_inj(ps, prm["kdn"], prm["kup"], prm["fp0"], 1, nprocs)
    fz = O.const_inj(vz, prm["kdn"], prm["kup"], prm["fz0"], 0, nprocs)
    time, rows, mm = 0.0, [], []
    for _ in range(prm["step"]):
        dt = O.CFL_condition(prm["CFL"], ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
        cc = O.cond_check(ps, vz, fp, fz, time, p, prm["kup"], nprocs)
        rows.append(cc["energy_bal"])
        mm.append(cc["max_min"])
        time += dt
        # last-stage forcing of this step, as the balance of the next iteration sees it
        c1, c2 = ps, vz
        ps, vz = O.rk_stages_const_inj(ps, vz, dt, p, prm["kdn"], prm["kup"], prm["fp0"], prm["fz0"],
                                       nprocs=nprocs)
        fp, fz = _last_stage_forcing(c1, c2, dt, p, prm, nprocs)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(gps.cpu().numpy(), ps) < 1e-11
    bal = np.loadtxt(str(tmp_path / "energy_bal.txt"))
    assert np.allclose(bal, np.array(rows), rtol=1e-9, atol=1e-12)
    got_mm = np.loadtxt(str(tmp_path / "max_min.txt"))
    assert np.allclose(got_mm, np.array(mm), rtol=1e-9, atol=1e-10)


def _last_stage_forcing(ps, vz, dt, p, prm, nprocs=1):
    """fp, fz computed at the top of stage o = 1 (from that stage's input)."""
    g = O.grid(ps.shape[1])
    upd = (g.ka2 <= g.kmax) & (g.ka2 > g.tiny)
    ka2 = np.where(upd, g.ka2, 1.0)
    c1, c2 = ps.copy(), vz.copy()
    for o in range(4, 0, -1):
        fp = O.const_inj(c1, prm["kdn"], prm["kup"], prm["fp0"], 1, nprocs)
        fz = O.const_inj(c2, prm["kdn"], prm["kup"], prm["fz0"], 0, nprocs)
        if o == 1:
            return fp, fz
        c4 = O.laplak2(c1)
        c5 = O.poisson(c1, c4)
        c6 = O.poisson(c1, c2)
        c3 = O.derivk2(c1, 2)
        c4 = O.derivk2(c2, 2)
        tmp1, tmp2 = dt / float(o), 1.0 / ka2
        t3 = 1.0 + (p.nu * ka2 ** p.nn + p.hnu * tmp2 ** p.mm) * tmp1
        n1 = (ps + ((-c5 + 2 * p.omega * c4) * tmp2 + fp) * tmp1) / t3
        t3 = 1.0 + (p.nuv * ka2 ** p.nnv + p.hnuv * tmp2 ** p.mmv) * tmp1
        n2 = (vz + (c6 + 2 * p.omega * c3 + fz) * tmp1) / t3
        c1, c2 = np.where(upd, n1, 0.0), np.where(upd, n2, 0.0)


def test_driver_balance_files_close_their_budgets(tmp_path, cuda):
    """The acceptance check of the reference (BalanceTest_Rot2D3C.py, restated
    in rot2d3c_amd.balance) on the files the GPU driver writes: with random
    forcing and Omega != 0 the total energy and helicity budgets close within
    4 standard errors; with Omega = 0 so do the 2-D, vz and enstrophy ones.
    (The oracle's loop gives at most 2.8 standard errors on these inputs.)"""
    from rot2d3c_amd import balance as B, driver, io as rio
    n = 32
    for omega in (0.0, 1.5):
        print("driver budgets n=32 omega=%g" % omega, flush=True)
        d = tmp_path / ("om%g" % omega)
        d.mkdir()
        prm = dict(PRM, step=500, cstep=1, sstep=10 ** 6, tstep=10 ** 6, omega=omega, CFL=0.3,
                   nu=1e-2, nuv=1e-2, fp0=1.0, fz0=1.0, kdn=1.0, kup=4.0)
        rio.write_parameter_inp(str(d / "parameter.inp"), prm)
        rio.write_status_inp(str(d / "status.inp"), 0, 0.0)
        driver.Run(n, str(d), "cuda:0").run()
        bud = B.budgets_from_files(str(d))
        closing = ["energy", "helicity"] + (["energy_2d", "energy_z", "enstrophy"] if omega == 0 else [])
        for k in closing:
            m, se = bud[k]
            assert abs(m) < 4 * se, (omega, k, m, se)
