"""GPU parity: the HIP kernels (through the C ABI) against the numpy oracle on
the same seeded inputs.  FP64 tolerance: relative L2 <= 1e-12 for one
operator application and for 10 time steps (the bar stated for this path)."""
import math

import numpy as np
import pytest

from conftest import random_state, rel_l2
from oracle import spectral as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
SIZES = [16, 32, 64, 256]


def _S():
    from rot2d3c_amd import spectral as S
    return S


def _dev(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def _host(t):
    return t.detach().cpu().numpy()


def _rand_spec(n, seed, hermitian=True):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n // 2 + 1, n)) + 1j * rng.standard_normal((n // 2 + 1, n))
    return a * n


@pytest.mark.parametrize("n", SIZES + [512, 2048, 4096])
def test_c2r_matches_oracle(n, cuda):
    print("c2r n=%d" % n, flush=True)
    S = _S()
    a = _rand_spec(n, 1)  # non-Hermitian kx=0 column and complex DC/Nyquist: c2r projects
    r = _host(S.fftp2d_complex_to_real(_dev(a, cuda)))
    assert rel_l2(r, O.fftp2d_complex_to_real(a)) < TOL


@pytest.mark.parametrize("n", SIZES + [512, 2048, 4096])
def test_r2c_matches_oracle(n, cuda):
    print("r2c n=%d" % n, flush=True)
    S = _S()
    r = np.random.default_rng(2).standard_normal((n, n))
    c = _host(S.fftp2d_real_to_complex(_dev(r, cuda)))
    assert rel_l2(c, O.fftp2d_real_to_complex(r)) < TOL


@pytest.mark.parametrize("n", [16, 64, 256])
def test_spectral_operators_exact(n, cuda):
    print("derivk2/laplak2/inv_laplak2 n=%d" % n, flush=True)
    S = _S()
    a = _rand_spec(n, 3)
    ta = _dev(a, cuda)
    for d in (0, 1, 2):
        assert rel_l2(_host(S.derivk2(ta, d)), O.derivk2(a, d)) < 1e-15
    assert rel_l2(_host(S.laplak2(ta)), O.laplak2(a)) < 1e-15
    assert rel_l2(_host(S.inv_laplak2(ta)), O.inv_laplak2(a)) < 1e-15


@pytest.mark.parametrize("n", SIZES + [4096])
def test_poisson_matches_oracle(n, cuda):
    print("poisson n=%d" % n, flush=True)
    S = _S()
    a = _rand_spec(n, 4)
    b = _rand_spec(n, 5)
    c = _host(S.poisson(_dev(a, cuda), _dev(b, cuda)))
    assert rel_l2(c, O.poisson(a, b)) < TOL


@pytest.mark.parametrize("n", [16, 64, 256])
def test_energy_inerprod_match_oracle(n, cuda):
    print("energy/inerprod n=%d" % n, flush=True)
    S = _S()
    a = _rand_spec(n, 6)
    b = _rand_spec(n, 7)
    ta, tb = _dev(a, cuda), _dev(b, cuda)
    for kin in (-2, -1, 0, 1, 2, 3):
        e = S.energy(ta, kin)
        assert abs(e - O.energy(a, kin)) <= 1e-13 * abs(O.energy(a, kin))
        ip = S.inerprod(ta, tb, kin)
        ref = O.inerprod(a, b, kin)
        assert abs(ip - ref) <= 1e-12 * max(abs(ref), O.energy(a, kin) ** 0.5 * O.energy(b, kin) ** 0.5)


@pytest.mark.parametrize("n", [32, 256])
def test_cfl_matches_oracle(n, cuda):
    print("cfl n=%d" % n, flush=True)
    S = _S()
    ps, vz = random_state(n, 8)
    dt = S.CFL_condition(0.5, _dev(ps, cuda), _dev(vz, cuda), 1e-3, 1, 2e-3, 2, 1.5)
    ref = O.CFL_condition(0.5, ps, vz, 1e-3, 1, 2e-3, 2, 1.5)
    assert abs(dt - ref) <= 1e-13 * ref


@pytest.mark.parametrize("n", [32, 64])
def test_forcing_fill_matches_oracle(n, cuda):
    print("forcing fill n=%d" % n, flush=True)
    S = _S()
    kup = 5.0
    seed = 11
    for _ in range(40):  # many draws: every placement branch (kx=0 column, ky<0, no forcing)
        fp_ref, seed_ref = O.rand_force(kup, kup, 1.0, 2e-3, seed, 1, n)
        fp, seed_new = S.rand_force(kup, kup, 1.0, 2e-3, seed, 1, n, cuda)
        assert seed_new == seed_ref
        assert np.array_equal(_host(fp), fp_ref)
        seed = seed_new


def _run_both(n, steps, p, dtmode, seed=9, kup=4.0, cuda=None):
    S = _S()
    ps, vz = random_state(n, seed)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    fixed = dtmode if isinstance(dtmode, float) else None
    cfl = 0.5
    ops, ovz, _, odts, _, _ = O.run_random_forcing(ps.copy(), vz.copy(), steps, cfl, p, kup, kup,
                                                   1.0, 1.0, seed, fixed_dt=fixed)
    import torch
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    dto = torch.zeros(1, dtype=torch.float64, device=cuda)
    s1 = seed + 1
    s0 = seed
    dt = fixed if fixed is not None else S.CFL_condition(cfl, tps, tvz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
    # pre-loop rand_force pair (advances both seeds, src/Rot2D3C.para.f90:377-380)
    _, _, _, s0 = S.rand_force_draw(kup, s0)
    _, _, _, s1 = S.rand_force_draw(kup, s1)
    dts = []
    for _ in range(steps):
        kx, ky, ph, s0 = S.rand_force_draw(kup, s0)
        fpd = S.ForceDraw(kx, ky, ph, 1.0, kup)
        kx, ky, ph, s1 = S.rand_force_draw(kup, s1)
        fzd = S.ForceDraw(kx, ky, ph, 1.0, 1.0)
        if fixed is not None:
            S.rk4_step(tps, tvz, sp, fpd, fzd, dt=fixed, dt_out=dto)
        else:
            S.rk4_step(tps, tvz, sp, fpd, fzd, cfl=cfl, dt_out=dto)
        dts.append(float(dto.item()))
    return _host(tps), _host(tvz), ops, ovz, np.array(dts), np.array(odts)


@pytest.mark.parametrize("n", SIZES + [512, 1024])
def test_step_cfl_10_steps(n, cuda):
    print("10 CFL steps n=%d" % n, flush=True)
    p = O.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 2, 1, 2.0)
    gps, gvz, ops, ovz, gdt, odt = _run_both(n, 10, p, "cfl", cuda=cuda)
    assert np.max(np.abs(gdt - odt) / odt) < 1e-12
    assert rel_l2(gps, ops) < TOL
    assert rel_l2(gvz, ovz) < TOL


@pytest.mark.parametrize("n", [64])
def test_step_fixed_dt_10_steps(n, cuda):
    print("10 fixed-dt steps n=%d" % n, flush=True)
    p = O.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    gps, gvz, ops, ovz, gdt, odt = _run_both(n, 10, p, 2e-3, cuda=cuda)
    assert np.all(gdt == 2e-3)
    assert rel_l2(gps, ops) < TOL
    assert rel_l2(gvz, ovz) < TOL


@pytest.mark.parametrize("n,mode", [(16, "cfl"), (256, "cfl"), (512, 1e-3), (2048, "cfl")])
def test_run_equals_repeated_steps_bitwise(n, mode, cuda):
    """r2d3c_run (chained steps) == n calls of r2d3c_step, bit for bit, from an
    initial state that is NOT dealiased (exercises the first step's full
    write and zeroing of kx > n/3)."""
    print("run vs steps n=%d mode=%s" % (n, mode), flush=True)
    import torch
    S = _S()
    rng = np.random.default_rng(21)
    ps = O.fftp2d_real_to_complex(rng.standard_normal((n, n)))
    vz = O.fftp2d_real_to_complex(rng.standard_normal((n, n)))
    ps *= 1.0 / np.sqrt(O.energy(ps, 1))
    vz *= 1.0 / np.sqrt(O.energy(vz, 0))
    p = S.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 2, 1, 2.0)
    steps, kup = 5, 4.0
    from rot2d3c_amd import ic
    sched = ic.forcing_schedule(kup, 1.0, 1.0, 3, steps, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    kw = dict(dt=mode) if isinstance(mode, float) else dict(cfl=0.5)
    a_ps, a_vz = _dev(ps, cuda), _dev(vz, cuda)
    b_ps, b_vz = _dev(ps, cuda), _dev(vz, cuda)
    dts_a = torch.zeros(steps, dtype=torch.float64, device=cuda)
    dts_b = torch.zeros(steps, dtype=torch.float64, device=cuda)
    for k in range(steps):
        S.rk4_step(a_ps, a_vz, p, draws[k][0], draws[k][1], dt_out=dts_a[k:k + 1], **kw)
    S.rk4_run(b_ps, b_vz, p, draws, dt_out=dts_b, **kw)
    assert torch.equal(dts_a, dts_b)
    assert torch.equal(torch.view_as_real(a_ps), torch.view_as_real(b_ps))
    assert torch.equal(torch.view_as_real(a_vz), torch.view_as_real(b_vz))
    g = O.grid(n)
    assert np.all(_host(b_ps)[~((g.ka2 <= g.kmax) & (g.ka2 > g.tiny))] == 0)


@pytest.mark.parametrize("n", [32, 256])
def test_cond_check_matches_oracle(n, cuda):
    print("cond_check n=%d" % n, flush=True)
    S = _S()
    ps, vz = random_state(n, 31)
    kup = 5.0
    fp, _ = O.rand_force(kup, kup, 1.0, 1e-3, 3, 1, n)
    fz, _ = O.rand_force(kup, kup, 0.5, 1e-3, 8, 0, n)
    p = O.Params(1e-3, 0.1, 1, 2, 2e-3, 0.2, 2, 1, 1.5)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    for nprocs in (1, 4):  # extrema: MPI_MAX of per-slab values with 4 ranks
        ref = O.cond_check(ps, vz, fp, fz, 0.125, p, kup, nprocs)
        got = S.cond_check(_dev(ps, cuda), _dev(vz, cuda), _dev(fp, cuda), _dev(fz, cuda), 0.125,
                           sp, kup, nprocs)
        for name in ref:
            r, g = np.array(ref[name]), np.array(got[name])
            scale = np.maximum(np.abs(r), 1e-300)
            assert np.all(np.abs(g - r) <= 1e-12 * scale + 1e-14 * np.abs(r).max()), (name, g, r)


@pytest.mark.parametrize("n", [32, 256, 2048])
def test_spectrum_transfers_match_oracle(n, cuda):
    print("spectrum/transfers n=%d" % n, flush=True)
    S = _S()
    ps, vz = random_state(n, 41)
    ek, ezk = S.spectrum(_dev(ps, cuda), _dev(vz, cuda))
    oek, oezk = O.spectrum(ps, vz)
    assert rel_l2(ek, oek) < 1e-13 and rel_l2(ezk, oezk) < 1e-13
    (f0, f1, f2), (x0, x1, x2) = S.transfers(_dev(ps, cuda), _dev(vz, cuda))
    (of0, of1, of2), (ox0, ox1, ox2) = O.transfers(ps, vz)
    for a, b in ((f0, of0), (f1, of1), (f2, of2)):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    for a, b in ((x0, ox0), (x1, ox1), (x2, ox2)):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))


@pytest.mark.parametrize("n,nprocs", [(32, 1), (256, 1), (32, 2), (64, 4), (256, 8)])
def test_const_inj_matches_oracle(n, nprocs, cuda):
    """nprocs > 1: the random numbers of a P-rank run (seed = rank per slab)."""
    print("const_inj n=%d nprocs=%d" % (n, nprocs), flush=True)
    S = _S()
    ps, vz = random_state(n, 51)
    # the wide band reaches past rank 0's kx slab, the narrow one does not
    for kdn, kup in ((3.0, 7.0), (3.0, n / 3.0)):
        for kin, f0 in ((1, 0.7), (0, 1.1)):
            got = _host(S.const_inj(_dev(ps, cuda), kdn, kup, f0, kin, nprocs))
            ref = O.const_inj(ps, kdn, kup, f0, kin, nprocs)
            assert rel_l2(got, ref) < 1e-13
    if nprocs > 1:  # the other ranks' random numbers differ from one rank's
        one = O.const_inj(ps, 3.0, n / 3.0, 0.7, 1, 1)
        assert rel_l2(one, O.const_inj(ps, 3.0, n / 3.0, 0.7, 1, nprocs)) > 1e-3


@pytest.mark.parametrize("n,nprocs", [(32, 1), (64, 1), (64, 2)])
def test_step_iflow2_10_steps(n, nprocs, cuda):
    """iflow = 2: forcing from each stage's own state (stage-by-stage path)."""
    print("10 iflow=2 steps n=%d nprocs=%d" % (n, nprocs), flush=True)
    import torch
    S = _S()
    ps, vz = random_state(n, 61)
    p = O.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 1, 2, 1.0)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    # with 2 ranks the band reaches into rank 1's kx slab (kx >= 17)
    ci = (2.0, 5.0 if nprocs == 1 else 20.0, 0.5, 0.3, nprocs)
    ops, ovz = ps.copy(), vz.copy()
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    dto = torch.zeros(1, dtype=torch.float64, device=cuda)
    for _ in range(10):
        dt = O.CFL_condition(0.5, ops, ovz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
        ops, ovz = O.rk_stages_const_inj(ops, ovz, dt, p, *ci[:4], nprocs=nprocs)
        S.rk4_step(tps, tvz, sp, cfl=0.5, dt_out=dto, const_inj=ci)
        assert abs(float(dto.item()) - dt) <= 1e-12 * dt
    assert rel_l2(_host(tps), ops) < TOL
    assert rel_l2(_host(tvz), ovz) < TOL


def test_bench_workload_2048_10_steps(cuda):
    """The 2048^2 workload the bench times (nu = nuv = 5e-5, nu+ = 0.5,
    Omega = 5, random forcing on kf = 40 for psi and vz at injection 1.0,
    fixed dt = 2e-4, seeded random-phase start), 10 chained steps through
    r2d3c_run with the plan's default row schedule, against the oracle's
    iflow = 3 loop on the same input: relative L2 <= 1e-12."""
    print("bench workload 2048, 10 steps", flush=True)
    import torch
    from rot2d3c_amd import ic
    S = _S()
    n, steps, dt, kf, seed = 2048, 10, 2e-4, 40.0, 0
    p = O.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    ps, vz = ic.random_phase_ic(n, seed=1234)
    sched = ic.forcing_schedule(kf, 1.0, 1.0, seed, steps, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    S.rk4_run(tps, tvz, sp, draws, dt=dt)
    torch.cuda.synchronize()
    ops, ovz, *_ = O.run_random_forcing(ps.copy(), vz.copy(), steps, 0.0, p, kf, kf, 1.0, 1.0, seed,
                                        fixed_dt=dt)
    assert rel_l2(_host(tps), ops) < TOL
    assert rel_l2(_host(tvz), ovz) < TOL


def test_largest_fused_size_4096_one_step(cuda):
    """n = 4096, the largest size of the fused step (its row kernels run at
    the register limit of 1024-thread blocks): one fixed-dt step with
    random forcing through r2d3c_run against the oracle's iflow = 3 loop,
    relative L2 <= 1e-12."""
    print("4096, 1 step", flush=True)
    import torch
    from rot2d3c_amd import ic
    S = _S()
    n, steps, dt, kf, seed = 4096, 1, 1e-4, 64.0, 5
    p = O.Params(2e-5, 0.5, 1, 2, 2e-5, 0.5, 1, 2, 3.0)
    ps, vz = ic.random_phase_ic(n, seed=77)
    sched = ic.forcing_schedule(kf, 1.0, 1.0, seed, steps, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    S.rk4_run(tps, tvz, sp, draws, dt=dt)
    torch.cuda.synchronize()
    ops, ovz, *_ = O.run_random_forcing(ps.copy(), vz.copy(), steps, 0.0, p, kf, kf, 1.0, 1.0, seed,
                                        fixed_dt=dt)
    assert rel_l2(_host(tps), ops) < TOL
    assert rel_l2(_host(tvz), ovz) < TOL


def test_bench_workload_2048_10_cfl_steps(cuda):
    """The 2048^2 workload with the reference's CFL-adaptive dt (CFL = 0.27
    gives dt near 2e-4), 10 chained steps through r2d3c_run against the
    oracle's iflow = 3 loop: every step's dt to 1e-12, fields to 1e-12."""
    print("bench workload 2048, 10 CFL steps", flush=True)
    import torch
    from rot2d3c_amd import ic
    S = _S()
    n, steps, cfl, kf, seed = 2048, 10, 0.27, 40.0, 0
    p = O.Params(5e-5, 0.5, 1, 2, 5e-5, 0.5, 1, 2, 5.0)
    ps, vz = ic.random_phase_ic(n, seed=1234)
    sched = ic.forcing_schedule(kf, 1.0, 1.0, seed, steps, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    dts = torch.zeros(steps, dtype=torch.float64, device=cuda)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    S.rk4_run(tps, tvz, sp, draws, cfl=cfl, dt_out=dts)
    torch.cuda.synchronize()
    O.set_fft_workers(8)  # bitwise the same transforms, threaded
    try:
        ops, ovz, _, odts, _, _ = O.run_random_forcing(ps.copy(), vz.copy(), steps, cfl, p, kf,
                                                       kf, 1.0, 1.0, seed)
    finally:
        O.set_fft_workers(None)
    gdt = _host(dts)
    assert np.max(np.abs(gdt - np.array(odts)) / np.array(odts)) < 1e-12
    assert rel_l2(_host(tps), ops) < TOL
    assert rel_l2(_host(tvz), ovz) < TOL


@pytest.mark.parametrize("n,mode", [(32, "cfl"), (256, 1e-3), (1024, "cfl")])
def test_step_dense_forcing_matches_oracle(n, mode, cuda):
    """The dense forcing arrays of r2d3c_step (iflow = 1 and the closure's
    forcing), together with a rand_force draw, against the oracle's stages
    with fp = dense + the draw's array: three chained steps.  The dense
    arrays are not band-limited, so the update mask's handling of them is
    exercised too."""
    print("dense forcing n=%d mode=%s" % (n, mode), flush=True)
    import torch
    S = _S()
    rng = np.random.default_rng(71)
    ps, vz = random_state(n, 72)
    fpd = O.fftp2d_real_to_complex(rng.standard_normal((n, n))) * (3.0 / n)
    fzd = O.fftp2d_real_to_complex(rng.standard_normal((n, n))) * (2.0 / n)
    p = O.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 2, 1, 1.5)
    sp = S.Params(p.nu, p.hnu, p.nn, p.mm, p.nuv, p.hnuv, p.nnv, p.mmv, p.omega)
    tps, tvz = _dev(ps, cuda), _dev(vz, cuda)
    tfp, tfz = _dev(fpd, cuda), _dev(fzd, cuda)
    dto = torch.zeros(1, dtype=torch.float64, device=cuda)
    kup, seed = n / 8.0, 19
    ops, ovz = ps.copy(), vz.copy()
    for _ in range(3):
        if mode == "cfl":
            dt = O.CFL_condition(0.5, ops, ovz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
        else:
            dt = mode
        kx, ky, ph, s_next = S.rand_force_draw(kup, seed)
        fdraw, _ = O.rand_force(kup, kup, 0.7, dt, seed, 1, n)
        seed = s_next
        ops, ovz = O.rk_stages(ops, ovz, fpd + fdraw, fzd, dt, p)
        kw = dict(cfl=0.5) if mode == "cfl" else dict(dt=mode)
        S.rk4_step(tps, tvz, sp, S.ForceDraw(kx, ky, ph, 0.7, kup), None, dt_out=dto,
                   dense=(tfp, tfz), **kw)
        assert abs(float(dto.item()) - dt) <= 1e-12 * dt
    assert rel_l2(_host(tps), ops) < TOL
    assert rel_l2(_host(tvz), ovz) < TOL


def test_run_rejects_steps_with_different_iflow2_settings(cuda):
    """r2d3c_run builds the iflow = 2 random-number table from the first
    step's band, so steps with another band or forcing kind are rejected
    (before any device work) instead of being applied with the wrong table."""
    print("r2d3c_run mixed iflow=2 settings", flush=True)
    import ctypes
    import torch
    from rot2d3c_amd import _lib
    S = _S()
    n = 32
    pl = S.Plan.get(n, cuda)
    ps = torch.zeros((n // 2 + 1, n), dtype=torch.complex128, device=cuda)
    vz = torch.zeros_like(ps)
    cp = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 1.0).c()
    for second in (dict(ci_enabled=1, ci_kdn=2.0, ci_kup=6.0), dict(ci_enabled=0)):
        arr = (_lib.Forcing * 2)()
        arr[0].ci_enabled, arr[0].ci_kdn, arr[0].ci_kup = 1, 2.0, 5.0
        for k, v in second.items():
            setattr(arr[1], k, v)
        rc = _lib.lib().r2d3c_run(pl.handle, ctypes.byref(cp), ctypes.c_void_p(ps.data_ptr()),
                                  ctypes.c_void_p(vz.data_ptr()), arr, 2, _lib.R2D3C_DT_FIXED,
                                  1e-3, None, None)
        assert rc == -1
        assert b"same ci_enabled" in _lib.lib().r2d3c_last_error()
