"""GPU: ensembles of independent runs (rot2d3c_amd.ensemble) and row-kernel
schedules.  Every run of an ensemble must equal the same run advanced alone
(r2d3c_run, fixed dt, same forcing, same schedule) bit for bit."""
import ctypes
import json

import numpy as np
import pytest
import torch

from conftest import random_state, rel_l2
from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _spec(sizes, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for r, n in enumerate(sizes):
        out.append(dict(n=n, seed=10 + r, omega=float(rng.uniform(0, 10)), nu=float(rng.uniform(1e-3, 3e-3)),
                        hnu=float(rng.uniform(0.05, 0.2)), kf=max(2.0, n / 16.0), f0=1.0,
                        dt=float(rng.uniform(0.5, 1.0)) * 1e-3 * 64.0 / n))
    return out


def _members(spec, cuda):
    from rot2d3c_amd import ensemble as E, spectral as S
    ms, host = [], []
    for r in spec:
        ps, vz = random_state(r["n"], r["seed"])
        host.append((ps, vz))
        p = S.Params(r["nu"], r["hnu"], 1, 2, r["nu"] * 1.5, r["hnu"] * 0.5, 1, 2, r["omega"])
        ms.append(E.Member(torch.from_numpy(ps).to(cuda), torch.from_numpy(vz).to(cuda), p, r["dt"]))
    return ms, host


def _solo(m, host, draws, nsteps, cuda):
    from rot2d3c_amd import spectral as S
    ps = torch.from_numpy(host[0]).to(cuda)
    vz = torch.from_numpy(host[1]).to(cuda)
    S.rk4_run(ps, vz, m.params, draws, nsteps, dt=m.dt)
    return ps, vz


@pytest.mark.parametrize("nsteps", [1, 6, 20])
def test_ensemble_runs_equal_solo_runs_bitwise(nsteps, cuda):
    """Mixed sizes (three classes, interleaved), per-run Omega, nu, dt and
    forcing; one step (no graph) and several (captured step graph)."""
    print("ensemble vs solo, %d steps" % nsteps, flush=True)
    from rot2d3c_amd import ensemble as E
    spec = _spec([32, 64, 32, 256, 64, 32])
    ms, host = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    table = E.forcing_table(spec, nsteps)
    ens.run(nsteps, table)
    torch.cuda.synchronize()
    for r, (m, h) in enumerate(zip(ms, host)):
        draws = [table[st][r] for st in range(nsteps)]
        ps, vz = _solo(m, h, draws, nsteps, cuda)
        assert torch.equal(m.ps, ps) and torch.equal(m.vz, vz), "run %d (n=%d)" % (r, m.n)
    ens.close()


def test_ensemble_second_run_call_and_no_forcing(cuda):
    """A second run() reuses the captured graph with a new forcing table; a
    third runs without forcing (a new graph): still equal to solo runs."""
    print("ensemble repeated calls", flush=True)
    from rot2d3c_amd import ensemble as E
    spec = _spec([64, 128, 64], seed=5)
    ms, host = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    t1 = E.forcing_table(spec, 4)
    t2 = E.forcing_table([dict(r, seed=r["seed"] + 100) for r in spec], 3)
    ens.run(4, t1)
    ens.run(3, t2)
    ens.run(2, None)
    torch.cuda.synchronize()
    for r, (m, h) in enumerate(zip(ms, host)):
        ps, vz = _solo(m, h, [t1[st][r] for st in range(4)], 4, cuda)
        from rot2d3c_amd import spectral as S
        S.rk4_run(ps, vz, m.params, [t2[st][r] for st in range(3)], 3, dt=m.dt)
        S.rk4_run(ps, vz, m.params, None, 2, dt=m.dt)
        torch.cuda.synchronize()
        assert torch.equal(m.ps, ps) and torch.equal(m.vz, vz), "run %d" % r
    ens.close()


def test_ensemble_run_matches_oracle(cuda):
    """One ensemble run against the oracle's loop (fixed dt, rand_force)."""
    print("ensemble vs oracle", flush=True)
    from rot2d3c_amd import ensemble as E
    spec = _spec([32, 64], seed=9)
    ms, host = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    nsteps = 8
    table = E.forcing_table(spec, nsteps)
    ens.run(nsteps, table)
    torch.cuda.synchronize()
    for r, (m, (ps, vz)) in enumerate(zip(ms, host)):
        q = m.params
        p = O.Params(q.nu, q.hnu, q.nn, q.mm, q.nuv, q.hnuv, q.nnv, q.mmv, q.omega)
        # the oracle's iflow = 3 loop draws the same sequence from the run's seed
        ps, vz, *_ = O.run_random_forcing(ps, vz, nsteps, 0.0, p, 1.0, spec[r]["kf"], spec[r]["f0"],
                                          spec[r]["f0"], spec[r]["seed"], fixed_dt=m.dt)
        assert rel_l2(m.ps.cpu().numpy(), ps) < 1e-12 and rel_l2(m.vz.cpu().numpy(), vz) < 1e-12
    ens.close()


@pytest.mark.parametrize("n", [256, 512, 1024, 2048])
def test_schedules_of_one_kernel_agree_bit_for_bit(n, cuda):
    """Schedules that run the same compiled row kernel (blocks per CU of the
    persistent kernel at a fixed rows per block) change only the launch
    shape: bit-identical results, and a schedule repeats bit for bit.  The
    three row kernels (non-persistent; persistent with 2 or 1 rows per block)
    are compiled separately, and the compiler may fuse their multiply-adds
    differently: they agree within rounding (1e-12 relative L2, 3 steps)."""
    print("schedules n=%d" % n, flush=True)
    from rot2d3c_amd import spectral as S, _lib
    from rot2d3c_amd._lib import check
    ps0, vz0 = random_state(n, 77)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    pl = S.Plan.get(n, cuda)
    assert _lib.lib().r2d3c_plan_set_schedule(pl.handle, 1, 0) != 0  # bpc 0 takes 2 rows
    kernels = [[(2, 0), (2, 0)], [(2, 1), (2, 2), (2, 1)]] + ([[(1, 1), (1, 2)]] if n >= 512 else [])
    firsts = []
    for group in kernels:
        outs = []
        for rw, bpc in group:
            check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, rw, bpc), "set_schedule")
            ps = torch.from_numpy(ps0).to(cuda)
            vz = torch.from_numpy(vz0).to(cuda)
            S.rk4_run(ps, vz, p, None, 3, dt=2e-4)
            torch.cuda.synchronize()
            outs.append((ps, vz))
        for ps, vz in outs[1:]:
            assert torch.equal(ps, outs[0][0]) and torch.equal(vz, outs[0][1]), group
        firsts.append(outs[0])
    check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, 2, 0), "set_schedule")
    for ps, vz in firsts[1:]:
        assert rel_l2(ps.cpu().numpy(), firsts[0][0].cpu().numpy()) < 1e-12
        assert rel_l2(vz.cpu().numpy(), firsts[0][1].cpu().numpy()) < 1e-12


@pytest.mark.parametrize("n", [256, 512, 2048])
def test_column_radix_forms(n, cuda):
    """The column transforms' two radix schedules (four-step, radix-8
    Stockham) agree within rounding over 3 steps (1e-12 relative L2), each
    repeats bit for bit, below 512 they are the same code (bit-identical),
    and the 3-step result stays within 1e-12 of the oracle at 512."""
    print("column forms n=%d" % n, flush=True)
    from rot2d3c_amd import spectral as S, _lib
    from rot2d3c_amd._lib import check
    ps0, vz0 = random_state(n, 78)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    pl = S.Plan.get(n, cuda)
    assert _lib.lib().r2d3c_plan_set_column_radix(pl.handle, 2) != 0
    outs = {}
    try:
        for cs in (0, 1, 1):
            check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, cs), "set_column_radix")
            got = ctypes.c_int(-1)
            check(_lib.lib().r2d3c_plan_get_column_radix(pl.handle, ctypes.byref(got)), "get")
            assert got.value == cs
            ps = torch.from_numpy(ps0).to(cuda)
            vz = torch.from_numpy(vz0).to(cuda)
            S.rk4_run(ps, vz, p, None, 3, dt=2e-4)
            torch.cuda.synchronize()
            if cs in outs:
                assert torch.equal(ps, outs[cs][0]) and torch.equal(vz, outs[cs][1])
            outs[cs] = (ps, vz)
    finally:
        check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, 0), "set_column_radix")
    (a0, b0), (a1, b1) = outs[0], outs[1]
    if n < 512:
        assert torch.equal(a0, a1) and torch.equal(b0, b1)
    assert rel_l2(a1.cpu().numpy(), a0.cpu().numpy()) < 1e-12
    assert rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) < 1e-12
    if n == 512:
        q = O.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
        ps, vz = ps0.copy(), vz0.copy()
        for _ in range(3):
            ps, vz = O.rk_stages(ps, vz, 0 * ps, 0 * vz, 2e-4, q)
        assert rel_l2(a1.cpu().numpy(), ps) < 1e-12 and rel_l2(b1.cpu().numpy(), vz) < 1e-12


def test_ensemble_with_stockham_columns_equals_solo_runs(cuda):
    """An ensemble's runs under column form 1 equal their solo runs under
    the same schedule bit for bit."""
    print("ensemble column form 1", flush=True)
    from rot2d3c_amd import spectral as S, _lib, ensemble as E
    from rot2d3c_amd._lib import check
    n, nsteps = 512, 4
    ps0, vz0 = random_state(n, 79)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    ps_e, vz_e = torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda)
    ens = E.Ensemble([E.Member(ps_e, vz_e, p, 2e-4)], schedules={n: (2, 0, 1)})
    ens.run(nsteps, None)
    torch.cuda.synchronize()
    ens.close()
    pl = S.Plan.get(n, cuda)
    check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, 2, 0), "set_schedule")
    check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, 1), "set_column_radix")
    try:
        ps, vz = torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda)
        S.rk4_run(ps, vz, p, None, nsteps, dt=2e-4)
        torch.cuda.synchronize()
    finally:
        check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, 0), "set_column_radix")
    assert torch.equal(ps, ps_e) and torch.equal(vz, vz_e)


def test_ensemble_tune_keeps_state_and_runs_match_solo(cuda):
    """Ensemble.tune times schedules, row forms and stage launch forms inside
    the ensemble's step graph: it leaves every run's state as it was
    (bitwise), keeps valid schedules, and afterwards each run still equals its
    solo run under the chosen schedule and row form."""
    print("ensemble tune", flush=True)
    from rot2d3c_amd import ensemble as E, spectral as S, _lib, autotune
    from rot2d3c_amd._lib import check
    spec = _spec([256, 512, 256, 64])
    ms, host = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    before = [(m.ps.clone(), m.vz.clone()) for m in ms]
    sch, t = ens.tune(steps=3, trials=1)
    assert t > 0 and set(sch) == {64, 256, 512}
    for n, v in sch.items():
        assert tuple(v) in autotune.CANDIDATES and (v[2] == 0 or n >= 512)
    for m, (ps, vz) in zip(ms, before):
        assert torch.equal(m.ps, ps) and torch.equal(m.vz, vz)
    nsteps = 3
    ens.run(nsteps, None)
    torch.cuda.synchronize()
    try:
        for m, h in zip(ms, host):
            pl = S.Plan.get(m.n, cuda)
            rw, bpc, cs = sch[m.n]
            check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, rw, bpc), "set_schedule")
            check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, cs), "set_column_radix")
            check(_lib.lib().r2d3c_plan_set_row_split(pl.handle, ens.row_split[m.n]), "row_split")
            ps, vz = _solo(m, h, None, nsteps, cuda)
            torch.cuda.synchronize()
            assert torch.equal(ps, m.ps) and torch.equal(vz, m.vz), m.n
    finally:
        for n in sch:
            pl = S.Plan.get(n, cuda)
            check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, 2, 0), "set_schedule")
            check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, 0), "set_column_radix")
            check(_lib.lib().r2d3c_plan_set_row_split(pl.handle, -1), "row_split")
    ens.close()


@pytest.mark.parametrize("n", [64, 256, 1024, 2048])
def test_column_split_is_bitwise(n, cuda):
    """Split emission of the fused column kernel (each column's block
    repeated per emitted field) gives the same bits as the unsplit kernel:
    forced on, forced off and automatic, 4 CFL steps with random forcing
    through r2d3c_run, dt included."""
    print("column split n=%d" % n, flush=True)
    from rot2d3c_amd import spectral as S, _lib, ic
    from rot2d3c_amd._lib import check
    ps0, vz0 = random_state(n, 81)
    p = S.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 2, 1, 2.0)
    sched = ic.forcing_schedule(max(2.0, n / 16.0), 1.0, 1.0, 5, 4, S.rand_force_draw)
    draws = [(S.ForceDraw(*a), S.ForceDraw(*b)) for a, b in sched]
    pl = S.Plan.get(n, cuda)
    assert _lib.lib().r2d3c_plan_set_column_split(pl.handle, 2) != 0
    outs = []
    try:
        for mode in (0, 1, -1):
            check(_lib.lib().r2d3c_plan_set_column_split(pl.handle, mode), "set_column_split")
            got = ctypes.c_int(7)
            check(_lib.lib().r2d3c_plan_get_column_split(pl.handle, ctypes.byref(got)), "get")
            assert got.value == mode
            ps, vz = torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda)
            dts = torch.zeros(4, dtype=torch.float64, device=cuda)
            S.rk4_run(ps, vz, p, draws, cfl=0.5, dt_out=dts)
            torch.cuda.synchronize()
            outs.append((ps, vz, dts))
    finally:
        check(_lib.lib().r2d3c_plan_set_column_split(pl.handle, -1), "set_column_split")
    for ps, vz, dts in outs[1:]:
        assert torch.equal(ps, outs[0][0]) and torch.equal(vz, outs[0][1])
        assert torch.equal(dts, outs[0][2])


def test_ensemble_column_split_is_bitwise(cuda):
    """Ensemble runs with split emission forced on equal the unsplit ones."""
    print("ensemble column split", flush=True)
    from rot2d3c_amd import ensemble as E
    from rot2d3c_amd._lib import check
    spec = _spec([256, 512, 256])
    res = []
    for mode in (0, 1):
        ms, host = _members(spec, cuda)
        ens = E.Ensemble(ms, schedules=None)
        for n in ens.sizes:
            check(E._lib.lib().r2d3c_ensemble_set_column_split(ens.handle, n, mode), "split")
        table = E.forcing_table(spec, 5)
        ens.run(5, table)
        torch.cuda.synchronize()
        res.append([(m.ps.clone(), m.vz.clone()) for m in ms])
        ens.close()
    for (a, b), (c, d) in zip(res[0], res[1]):
        assert torch.equal(a, c) and torch.equal(b, d)


def test_autotune_picks_a_valid_schedule(cuda, tmp_path, monkeypatch):
    print("autotune n=512", flush=True)
    cache = tmp_path / "s.json"
    monkeypatch.setenv("R2D3C_SCHEDULE_CACHE", str(cache))
    from rot2d3c_amd import autotune
    key = "%s:512" % autotune.device_key(cuda)
    # an entry cached before the column form was tuned is tuned again
    cache.write_text(json.dumps({key: {"schedule": [1, 1], "ms_per_step": 1.0}}))
    sch = autotune.schedule(512, cuda, reps=2)
    assert sch[:2] in ((2, 0), (2, 1), (2, 2), (1, 1), (1, 2)) and sch[2] in (0, 1)
    assert autotune.schedule(512, cuda) == sch  # from the cache
    assert json.loads(cache.read_text())[key]["schedule"] == list(sch)


def _set_plan_schedule(n, sch, cuda):
    from rot2d3c_amd import spectral as S, _lib
    from rot2d3c_amd._lib import check
    pl = S.Plan.get(n, cuda)
    check(_lib.lib().r2d3c_plan_set_schedule(pl.handle, sch[0], sch[1]), "set_schedule")
    check(_lib.lib().r2d3c_plan_set_column_radix(pl.handle, sch[2]), "set_column_radix")


@pytest.mark.parametrize("schedules", [None, {256: (2, 0, 0), 512: (2, 0, 1), 1024: (2, 0, 1)}],
                         ids=["default", "stockham-columns"])
def test_config3_mix_runs_equal_solo_runs_and_oracle(schedules, cuda):
    """The batched workload itself: 24 x 256^2 + 12 x 512^2 + 4 x 1024^2 runs
    (own seed, Omega ~ U[0,10], nu = 1e-3*256/n, kf = n/32, fixed dt), 3 steps
    in one ensemble, with the default schedules and with the Stockham column
    form at 512 and 1024 (the schedules the ensemble tuner picks).  Every run
    equals its solo r2d3c_run under the same schedule bit for bit, and one run
    of each size (all four 1024^2 runs) matches the oracle's iflow = 3 loop
    at 1e-12."""
    from rot2d3c_amd import ensemble as E, spectral as S
    from rot2d3c_amd.ic import random_phase_ic
    nsteps = 3
    spec = E.mixed_ensemble_spec()
    print("config-3 mix, %d runs, %d steps, schedules=%s" % (len(spec), nsteps, schedules),
          flush=True)
    ens = E.build(spec, cuda, schedules=schedules)
    table = E.forcing_table(spec, nsteps)
    ens.run(nsteps, table)
    torch.cuda.synchronize()
    assert sorted(ens.sizes) == [256, 512, 1024]
    try:
        for n in ens.sizes:
            _set_plan_schedule(n, ens.schedules[n], cuda)
        for r, (m, sp) in enumerate(zip(ens.members, spec)):
            ps, vz = random_phase_ic(sp["n"], sp["seed"])
            sps, svz = _solo(m, (ps, vz), [table[st][r] for st in range(nsteps)], nsteps, cuda)
            assert torch.equal(m.ps, sps) and torch.equal(m.vz, svz), "run %d (n=%d)" % (r, m.n)
    finally:
        for n in ens.sizes:
            _set_plan_schedule(n, (2, 0, 0), cuda)
    checked = [0, 24] + [r for r, sp in enumerate(spec) if sp["n"] == 1024]
    O.set_fft_workers(8)
    try:
        for r in checked:
            sp, m = spec[r], ens.members[r]
            print("  oracle run %d n=%d" % (r, sp["n"]), flush=True)
            q = m.params
            p = O.Params(q.nu, q.hnu, q.nn, q.mm, q.nuv, q.hnuv, q.nnv, q.mmv, q.omega)
            ps, vz = random_phase_ic(sp["n"], sp["seed"])
            ops, ovz, *_ = O.run_random_forcing(ps, vz, nsteps, 0.0, p, sp["kf"], sp["kf"],
                                                sp["f0"], sp["f0"], sp["seed"], fixed_dt=m.dt)
            assert rel_l2(m.ps.cpu().numpy(), ops) < 1e-12, r
            assert rel_l2(m.vz.cpu().numpy(), ovz) < 1e-12, r
    finally:
        O.set_fft_workers(None)
    ens.close()


@pytest.mark.parametrize("n", [16, 64, 256, 512])
def test_row_split_forms_agree(n, cuda):
    """The row kernel with the three field pairs' inverse transforms in
    parallel (row split 1, the default for n <= 512) and the sequential one
    (0) agree within rounding over 3 CFL steps; with either form a chained
    run equals repeated single steps bit for bit."""
    print("row split forms n=%d" % n, flush=True)
    from rot2d3c_amd import spectral as S, _lib
    from rot2d3c_amd._lib import check
    pl = S.Plan.get(n, cuda)
    mode = ctypes.c_int()
    check(_lib.lib().r2d3c_plan_get_row_split(pl.handle, ctypes.byref(mode)), "get_row_split")
    assert mode.value == -1
    assert _lib.lib().r2d3c_plan_set_row_split(pl.handle, 2) != 0
    ps0, vz0 = random_state(n, 91)
    p = S.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 1, 2, 1.5)
    out = {}
    try:
        for m in (0, 1):
            check(_lib.lib().r2d3c_plan_set_row_split(pl.handle, m), "set_row_split")
            a = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
            b = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
            for _ in range(3):
                S.rk4_step(a[0], a[1], p, cfl=0.5)
            S.rk4_run(b[0], b[1], p, None, 3, cfl=0.5)
            torch.cuda.synchronize()
            assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), m
            out[m] = (a[0].cpu().numpy(), a[1].cpu().numpy())
    finally:
        check(_lib.lib().r2d3c_plan_set_row_split(pl.handle, -1), "set_row_split")
    assert rel_l2(out[1][0], out[0][0]) < 1e-12 and rel_l2(out[1][1], out[0][1]) < 1e-12


def test_ensemble_order_of_size_classes_does_not_change_results(cuda):
    """Size classes one after another, as concurrent graph branches and each
    on its own stream (the default) give the same bits, each run equal to
    its solo run (12 steps: the first, a replay of the graph of 8 steps, three
    replays of the one-step graph)."""
    print("ensemble class order", flush=True)
    from rot2d3c_amd import ensemble as E
    spec = _spec([32, 64, 32, 128, 64], seed=13)
    nsteps = 12
    table = E.forcing_table(spec, nsteps)
    outs = []
    for conc in (0, 1, 2):
        ms, host = _members(spec, cuda)
        ens = E.Ensemble(ms, schedules=None)
        assert ens.concurrent == 2
        ens.set_concurrent(conc)
        ens.run(nsteps, table)
        torch.cuda.synchronize()
        outs.append([(m.ps.clone(), m.vz.clone()) for m in ms])
        ens.close()
    for r, ((a0, b0), (a1, b1), (a2, b2)) in enumerate(zip(*outs)):
        assert torch.equal(a0, a1) and torch.equal(b0, b1), r
        assert torch.equal(a0, a2) and torch.equal(b0, b2), r
    ms, host = _members(spec, cuda)
    for r, (m, h) in enumerate(zip(ms, host)):
        ps, vz = _solo(m, h, [table[st][r] for st in range(nsteps)], nsteps, cuda)
        assert torch.equal(outs[0][r][0], ps) and torch.equal(outs[0][r][1], vz), r


@pytest.mark.parametrize("form,nsteps", [(1, 1), (1, 6), (2, 6)])
def test_all_size_stage_launches_equal_solo_runs_bitwise(form, nsteps, cuda):
    """One persistent launch per RK4 stage for all sizes (form 1), and one
    launch per phase for all sizes (form 2): with 64, 128, 256, 512 and 1024
    runs interleaved, with forcing, every run equals its solo r2d3c_run with
    these launches' schedule (sequential row kernel, Stockham columns) bit
    for bit; no grid barrier timed out."""
    print("all-size stage launch %d vs solo, %d steps" % (form, nsteps), flush=True)
    from rot2d3c_amd import ensemble as E, spectral as S, _lib
    spec = _spec([256, 64, 1024, 512, 128, 256, 64, 512], seed=21)
    ms, host = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    ens.set_stage_launch(form)
    if form == 1:
        assert ens.persistent_grid() >= 256 and ens.persistent_grid() % 16 == 0
    table = E.forcing_table(spec, nsteps)
    ens.run(nsteps, table)
    ens.sync()
    for r, (m, h) in enumerate(zip(ms, host)):
        pl = S.Plan.get(m.n, cuda).handle
        old = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
        _lib.check(_lib.lib().r2d3c_plan_get_schedule(pl, ctypes.byref(old[0]), ctypes.byref(old[1])), "get")
        _lib.check(_lib.lib().r2d3c_plan_get_row_split(pl, ctypes.byref(old[2])), "get")
        _lib.check(_lib.lib().r2d3c_plan_get_column_radix(pl, ctypes.byref(old[3])), "get")
        try:
            _lib.check(_lib.lib().r2d3c_plan_set_schedule(pl, 2, 0), "set")
            _lib.check(_lib.lib().r2d3c_plan_set_row_split(pl, 0), "set")
            _lib.check(_lib.lib().r2d3c_plan_set_column_radix(pl, 1), "set")
            draws = [table[st][r] for st in range(nsteps)]
            ps, vz = _solo(m, h, draws, nsteps, cuda)
            torch.cuda.synchronize()
        finally:
            _lib.lib().r2d3c_plan_set_schedule(pl, old[0].value, old[1].value)
            _lib.lib().r2d3c_plan_set_row_split(pl, old[2].value)
            _lib.lib().r2d3c_plan_set_column_radix(pl, old[3].value)
        assert torch.equal(m.ps, ps) and torch.equal(m.vz, vz), "run %d (n=%d)" % (r, m.n)
    ens.close()


def test_all_size_stage_launches_reject_unsupported_sizes(cuda):
    print("all-size stage launch: sizes", flush=True)
    from rot2d3c_amd import ensemble as E, _lib
    spec = _spec([32, 64], seed=4)
    ms, _ = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    for form in (1, 2):
        with pytest.raises(_lib.R2D3CError):
            ens.set_stage_launch(form)
    assert ens.persistent_grid() == 0 and ens.stage_launch == 0
    ens.close()
    # a closure cannot join an ensemble whose stages run as all-size launches
    from rot2d3c_amd import closure as L
    spec = _spec([64, 128], seed=6)
    ms, _ = _members(spec, cuda)
    ens = E.Ensemble(ms, schedules=None)
    ens.set_stage_launch(2)
    net = L.Net(1, 64, 64, cuda, seed=1)
    with pytest.raises(_lib.R2D3CError):
        ens.set_closure(0, net, [1.0, 1.0, 1.0], [1.0, 1.0])
    ens.set_stage_launch(0)
    ens.set_closure(0, net, [1.0, 1.0, 1.0], [1.0, 1.0])
    ens.close()
