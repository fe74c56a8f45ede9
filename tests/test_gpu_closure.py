"""GPU: the learned closure's kernels (rot2d3c_amd/csrc/closure.hip and the
closure helpers of rot2d3c.hip) against the numpy oracles."""
import math

import numpy as np
import pytest
import torch

from conftest import random_state, rel_l2
from oracle import cnn as C
from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _net(cuda, B, H, W, params):
    from rot2d3c_amd import closure as L
    return L.Net(B, H, W, cuda, params=params.astype(np.float32))


def test_param_layout_matches_oracle(cuda):
    from rot2d3c_amd import closure as L
    wo, bo, ci, co = L.layer_layout()
    offs, total = C.offsets()
    assert [o[0] for o in offs] == wo and [o[1] for o in offs] == bo
    assert ci == C.CI and co == C.CO
    from rot2d3c_amd import _lib
    assert _lib.lib().r2d3c_cnn_param_count() == total


def test_cnn_forward_matches_oracle(cuda):
    """BF16 MFMA forward against the oracle with the same bf16 rounding points;
    the remaining differences are FP32 vs FP64 accumulation."""
    print("cnn forward", flush=True)
    rng = np.random.default_rng(3)
    B, H, W = 2, 16, 64
    p = C.init_params(rng)
    x = rng.standard_normal((B, H, W, 3))
    net = _net(cuda, B, H, W, p)
    got = net.forward(torch.from_numpy(x.astype(np.float32)).to(cuda)).cpu().numpy()
    ref, _ = C.forward(p, x.astype(np.float32).astype(np.float64), emulate=True)
    assert rel_l2(got, ref) < 5e-3
    exact, _ = C.forward(p, x, emulate=False)
    assert rel_l2(got, exact) < 5e-2  # bf16 network vs the exact one
    net.close()


def test_cnn_gradients_match_oracle(cuda):
    """Loss and every layer's weight and bias gradient of the GPU backward
    pass against the oracle's (same bf16 rounding points)."""
    print("cnn gradients", flush=True)
    rng = np.random.default_rng(4)
    B, H, W = 2, 8, 64
    p = C.init_params(rng)
    x = rng.standard_normal((B, H, W, 3)).astype(np.float32)
    y = rng.standard_normal((B, H, W, 2)).astype(np.float32)
    net = _net(cuda, B, H, W, p)
    loss, g = net.gradients(torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda))
    g = g.cpu().numpy().astype(np.float64)
    rloss, rg = C.loss_and_grad(p, x.astype(np.float64), y.astype(np.float64), emulate=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    offs, _ = C.offsets()
    for l, (wo, bo) in enumerate(offs):
        assert rel_l2(g[wo:bo], rg[wo:bo]) < 2e-2, ("W", l)
        assert rel_l2(g[bo:bo + C.CO[l]], rg[bo:bo + C.CO[l]]) < 2e-2, ("b", l)
    # the padding's gradient is exactly zero
    W0 = g[offs[0][0]:offs[0][1]].reshape(9, 32, 16)
    assert np.all(W0[:, :, 3:] == 0)
    net.close()


def test_cnn_training_fits_a_teacher(cuda):
    """Adam on the GPU fits the outputs of a random teacher network."""
    print("cnn training", flush=True)
    rng = np.random.default_rng(5)
    B, H, W = 4, 16, 64
    teacher = C.init_params(rng, scale=0.5)
    x = rng.standard_normal((B, H, W, 3))
    y, _ = C.forward(teacher, x, emulate=False)
    net = _net(cuda, B, H, W, C.init_params(np.random.default_rng(6)))
    xt = torch.from_numpy(x.astype(np.float32)).to(cuda)
    yt = torch.from_numpy(y.astype(np.float32)).to(cuda)
    l0, _ = net.gradients(xt, yt)
    for _ in range(150):
        net.train_step(xt, yt, 2e-3)
    l1, _ = net.gradients(xt, yt)
    assert l1 < 0.2 * l0, (l0, l1)
    net.close()


def test_filter_of_a_single_mode(cuda):
    """G(cos(kx x + ky y)) on the coarse grid is sinc(kx h/2) sinc(ky h/2)
    cos(kx x + ky y) for a mode inside the coarse dealiased disc and 0 outside."""
    print("filter", flush=True)
    from rot2d3c_amd import closure as L
    nf, nc = 64, 16
    h = 2 * math.pi / nc
    xf = 2 * math.pi * np.arange(nf) / nf
    xc = 2 * math.pi * np.arange(nc) / nc
    sinc = lambda a: 1.0 if a == 0 else math.sin(a) / a
    for kx, ky, inside in ((3, -2, True), (0, 4, True), (6, 1, False)):
        R = np.cos(kx * xf[None, :] + ky * xf[:, None])
        a = torch.from_numpy(O.fftp2d_real_to_complex(R)).to(cuda)
        out = L.coarse_grain(a, nc, h).cpu().numpy()
        got = O.fftp2d_complex_to_real(out) / float(nc * nc)
        expect = (sinc(kx * h / 2) * sinc(ky * h / 2) if inside else 0.0) * \
            np.cos(kx * xc[None, :] + ky * xc[:, None])
        assert np.max(np.abs(got - expect)) < 1e-12, (kx, ky)


def test_closure_inputs_targets_and_forcing(cuda):
    print("closure inputs/targets/forcing", flush=True)
    from rot2d3c_amd import closure as L
    from rot2d3c_amd import spectral as S
    n = 32
    ps, vz = random_state(n, 7)
    g = O.grid(n)
    n2 = float(n * n)
    inp = L.network_inputs(torch.from_numpy(ps).to(cuda), torch.from_numpy(vz).to(cuda),
                           (2.0, 3.0, 5.0)).cpu().numpy()
    ref = [O.fftp2d_complex_to_real(O.laplak2(ps)) / n2 * 2.0, O.fftp2d_complex_to_real(vz) / n2 * 3.0,
           O.fftp2d_complex_to_real(ps) / n2 * 5.0]
    for ch in range(3):
        assert rel_l2(inp[..., ch], ref[ch]) < 1e-6
    rng = np.random.default_rng(8)
    f5, c5, f6, c6 = [random_state(n, 20 + k)[0] * rng.uniform(0.5, 2) for k in range(4)]
    upd = (g.ka2 <= g.kmax) & (g.ka2 > g.tiny)
    pl = S.Plan.get(n, cuda)
    dev = [torch.from_numpy(a).to(cuda) for a in (f5, c5, f6, c6)]
    tgt = torch.empty((n, n, 2), dtype=torch.float32, device=cuda)
    from rot2d3c_amd import _lib
    _lib.check(_lib.lib().r2d3c_closure_targets(pl.handle, *[S._ptr(t) for t in dev],
                                                L._dvec((1.0, 1.0)), 0.0, L._fptr(tgt),
                                                S._stream(cuda)),
               "targets")
    t5 = np.where(upd, f5 - c5, 0)
    t6 = np.where(upd, f6 - c6, 0)
    got = tgt.cpu().numpy()
    assert rel_l2(got[..., 0], O.fftp2d_complex_to_real(t5) / n2) < 1e-6
    assert rel_l2(got[..., 1], O.fftp2d_complex_to_real(t6) / n2) < 1e-6
    fp, fz = L.closure_forcing(tgt, n, (1.0, 1.0))
    ka2 = np.where(upd, g.ka2, 1.0)
    assert rel_l2(fp.cpu().numpy(), np.where(upd, -t5 / ka2, 0)) < 1e-6
    assert rel_l2(fz.cpu().numpy(), t6) < 1e-6
    # a low-k cutoff removes exactly the modes with |k| < kmin
    fp, fz = L.closure_forcing(tgt, n, (1.0, 1.0), kmin=4.0)
    hi = upd & (g.ka2 >= 16.0)
    assert rel_l2(fp.cpu().numpy(), np.where(hi, -t5 / ka2, 0)) < 1e-6
    assert np.all(fz.cpu().numpy()[~hi] == 0)


def test_closure_run_without_network_is_the_plain_solver(cuda):
    print("closure run fallback", flush=True)
    from rot2d3c_amd import closure as L
    from rot2d3c_amd import spectral as S
    n = 64
    ps, vz = random_state(n, 9)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 1.0)
    a = (torch.from_numpy(ps).to(cuda), torch.from_numpy(vz).to(cuda))
    b = (a[0].clone(), a[1].clone())
    run = L.ClosureRun(a[0], a[1], p)
    for _ in range(3):
        run.step(dt=1e-3)
        S.rk4_step(b[0], b[1], p, dt=1e-3)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def _fused_forcing(n, ps, vz, net, sin, sout, cuda):
    from rot2d3c_amd import _lib, closure as L, spectral as S
    pl = S.Plan.get(n, cuda)
    fp = torch.empty_like(ps)
    fz = torch.empty_like(ps)
    lib = _lib.lib()
    _lib.check(lib.r2d3c_closure_net_inputs(pl.handle, S._ptr(ps), S._ptr(vz), L._dvec(sin),
                                            net.handle, S._stream(cuda)), "net inputs")
    _lib.check(lib.r2d3c_cnn_forward_buffers(net.handle, 1, S._stream(cuda)), "forward")
    _lib.check(lib.r2d3c_closure_net_forcing(pl.handle, net.handle, L._dvec(sout), 0.0, S._ptr(fp),
                                             S._ptr(fz), S._stream(cuda)), "net forcing")
    return fp, fz


def test_fused_closure_io_matches_the_separate_steps(cuda):
    """closure_net_inputs -> forward_buffers -> closure_net_forcing equals
    closure_inputs -> cnn_forward -> closure_forcing up to rare bf16 rounding
    flips of the inputs (the two paths pack the row transforms differently)."""
    print("fused closure io", flush=True)
    from rot2d3c_amd import closure as L
    n = 64
    ps, vz = random_state(n, 31)
    ps, vz = torch.from_numpy(ps).to(cuda), torch.from_numpy(vz).to(cuda)
    net = L.Net(1, n, n, cuda, seed=3)
    sin, sout = (0.02, 0.5, 3.0), (7.0, 0.3)
    pred = net.forward(L.network_inputs(ps, vz, sin)[None].contiguous())
    fp0, fz0 = L.closure_forcing(pred[0], n, sout)
    fp1, fz1 = _fused_forcing(n, ps, vz, net, sin, sout, cuda)
    assert rel_l2(fp1.cpu().numpy(), fp0.cpu().numpy()) < 1e-2
    assert rel_l2(fz1.cpu().numpy(), fz0.cpu().numpy()) < 1e-2
    # with the same network input the output path is the same arithmetic
    fp2, fz2 = L.closure_forcing(torch.from_numpy(np.zeros((n, n, 2), np.float32)).to(cuda), n, sout)
    assert float(fp2.abs().max()) == 0.0 and float(fz2.abs().max()) == 0.0
    net.close()


def test_forcing_gain_scales_each_shell(cuda):
    """r2d3c_cnn_set_forcing_gain: the closure forcing made from the network's
    prediction is the one without the table times gain[int(|k| + 0.5)] (bit
    for bit: one multiplication per mode), shells past the table keep gain 1,
    and removing the table restores the original bits."""
    print("closure forcing gain", flush=True)
    from rot2d3c_amd import closure as L
    n = 64
    ps, vz = random_state(n, 33)
    ps, vz = torch.from_numpy(ps).to(cuda), torch.from_numpy(vz).to(cuda)
    net = L.Net(1, n, n, cuda, seed=5)
    sin, sout = (0.02, 0.5, 3.0), (7.0, 0.3)
    fp0, fz0 = _fused_forcing(n, ps, vz, net, sin, sout, cuda)
    gain = L.taper_gain(4.0, 9.0, 15) * 1.25
    net.set_forcing_gain(gain)
    fp1, fz1 = _fused_forcing(n, ps, vz, net, sin, sout, cuda)
    g = O.grid(n)
    shell = np.floor(np.sqrt(g.ka2) + 0.5).astype(int)
    w = np.where(shell < len(gain), gain[np.minimum(shell, len(gain) - 1)], 1.0)
    for a, b in ((fp0, fp1), (fz0, fz1)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.array_equal(b.real, a.real * w) and np.array_equal(b.imag, a.imag * w)
    net.set_forcing_gain(None)
    fp2, fz2 = _fused_forcing(n, ps, vz, net, sin, sout, cuda)
    assert torch.equal(fp2, fp0) and torch.equal(fz2, fz0)
    with pytest.raises(RuntimeError):
        net.set_forcing_gain(np.full(40, 1.0))  # more shells than 64/2 + 2
    net.close()


@pytest.mark.parametrize("n", [64, 256])
def test_ensemble_closure_run_tracks_the_stepwise_closure_run(cuda, n):
    """A closure run inside the ensemble's step graph against ClosureRun
    (separate kernels, same network and forcing): equal up to the inputs'
    bf16 rounding flips; a network of the wrong size is refused.  At 256 the
    ensemble's steady steps form the network's input inside its first layer
    and fold the forcing's passes into the first stage's launches."""
    print("ensemble closure run", n, flush=True)
    from rot2d3c_amd import closure as L, ensemble as E, spectral as S
    from rot2d3c_amd._lib import R2D3CError
    ps0, vz0 = random_state(n, 33)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    net_a = L.Net(1, n, n, cuda, seed=4)
    net_b = L.Net(1, n, n, cuda, seed=4)
    sin, sout = (0.02, 0.5, 3.0), (2.0, 0.2)
    ps = torch.from_numpy(ps0).to(cuda)
    vz = torch.from_numpy(vz0).to(cuda)
    ens = E.Ensemble([E.Member(ps, vz, p, 1e-3)], schedules=None)
    with pytest.raises(R2D3CError):
        ens.set_closure(0, L.Net(1, 2 * n, 2 * n, cuda), sin, sout)
    ens.set_closure(0, net_a, sin, sout)
    spec = [dict(kf=6.0, f0=1.0, seed=5)]
    table = E.forcing_table(spec, 5)
    ens.run(5, table)
    b = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
    run = L.ClosureRun(b[0], b[1], p, net_b, sin, sout, guard_every=0)
    for st in range(5):
        run.step(*table[st][0], dt=1e-3)
    torch.cuda.synchronize()
    assert rel_l2(ps.cpu().numpy(), b[0].cpu().numpy()) < 1e-6
    assert rel_l2(vz.cpu().numpy(), b[1].cpu().numpy()) < 1e-6
    # the closure changes the run: the plain solver differs
    c = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
    S.rk4_run(c[0], c[1], p, [table[st][0] for st in range(5)], 5, dt=1e-3)
    torch.cuda.synchronize()
    assert rel_l2(c[0].cpu().numpy(), b[0].cpu().numpy()) > 1e-9
    ens.close()


def test_backscatter_clip_properties(cuda):
    """closure.clip_forcing on random fields: modes that drained energy are
    untouched bit for bit; at the others the energy transfer becomes zero,
    the energy-neutral part Im(conj(psi) f) is kept and the removed part is
    parallel to the mode's coefficient; zero coefficients are left alone."""
    print("backscatter clip", flush=True)
    from rot2d3c_amd import closure as L
    n = 64
    rng = np.random.default_rng(41)
    shape = (n // 2 + 1, n)
    arr = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(4)]
    arr[0][3, :7] = 0.0  # modes with psi = 0
    arr[0][0, 0] = arr[1][0, 0] = 0.0  # the mean mode (|k| = 0 < kclip) is outside the clipping
    ps, vz, fp, fz = [torch.from_numpy(a).to(cuda) for a in arr]
    # on |k| >= 10 only: every mode below is untouched
    fp10, fz10 = fp.clone(), fz.clone()
    L.clip_forcing(ps, vz, fp10, fz10, kclip=10.0)
    kx = np.arange(n // 2 + 1)[:, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :]
    low = kx ** 2 + ky ** 2 < 100.0
    L.clip_forcing(ps, vz, fp, fz)
    torch.cuda.synchronize()
    for f10, f, f0 in ((fp10, fp, arr[2]), (fz10, fz, arr[3])):
        f10 = f10.cpu().numpy()
        assert np.array_equal(f10[low], f0[low]) and np.array_equal(f10[~low], f.cpu().numpy()[~low])
    for a, f0, f1 in ((arr[0], arr[2], fp.cpu().numpy()), (arr[1], arr[3], fz.cpu().numpy())):
        t0 = (np.conj(a) * f0).real
        t1 = (np.conj(a) * f1).real
        drain = t0 <= 0
        assert 0.3 < drain.mean() < 0.7
        assert np.array_equal(f1[drain], f0[drain])
        scale = np.abs(a) * np.abs(f0)
        assert np.all(np.abs(t1[~drain]) <= 1e-13 * scale[~drain])
        assert np.allclose((np.conj(a) * f1).imag, (np.conj(a) * f0).imag, rtol=1e-13, atol=1e-13)
        d = f0 - f1
        assert np.all(np.abs((np.conj(a) * d).imag) <= 1e-13 * scale + 1e-300)
    assert np.array_equal(fp.cpu().numpy()[3, :7], arr[2][3, :7])


def test_ensemble_clipped_closure_tracks_the_stepwise_run(cuda):
    """The clipped closure inside the ensemble's step graph against
    ClosureRun(kclip=10), and clipping changes the run."""
    print("ensemble clipped closure run", flush=True)
    from rot2d3c_amd import closure as L, ensemble as E, spectral as S
    n = 64
    ps0, vz0 = random_state(n, 35)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    sin, sout = (0.02, 0.5, 3.0), (2.0, 0.2)
    nets = [L.Net(1, n, n, cuda, seed=6) for _ in range(3)]
    spec = [dict(kf=6.0, f0=1.0, seed=7)]
    table = E.forcing_table(spec, 5)
    ps = torch.from_numpy(ps0).to(cuda)
    vz = torch.from_numpy(vz0).to(cuda)
    ens = E.Ensemble([E.Member(ps, vz, p, 1e-3)], schedules=None)
    ens.set_closure(0, nets[0], sin, sout, 0.0, kclip=10.0)
    ens.run(5, table)
    runs = []
    for net, kclip in ((nets[1], 10.0), (nets[2], 0.0)):
        b = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
        run = L.ClosureRun(b[0], b[1], p, net, sin, sout, guard_every=0, kclip=kclip)
        for st in range(5):
            run.step(*table[st][0], dt=1e-3)
        runs.append(b)
    torch.cuda.synchronize()
    assert rel_l2(ps.cpu().numpy(), runs[0][0].cpu().numpy()) < 1e-6
    assert rel_l2(vz.cpu().numpy(), runs[0][1].cpu().numpy()) < 1e-6
    assert rel_l2(runs[1][0].cpu().numpy(), runs[0][0].cpu().numpy()) > 1e-9
    ens.close()


def test_output_clamp_saturates_inference_only(cuda):
    """With an output clamp the forward pass returns exactly the clipped
    unclamped outputs (the saturation acts on the same fp32 values); with
    sign clipping exactly the outputs zeroed where they share the input's
    sign; gradients are those of the plain network."""
    print("cnn output clamp", flush=True)
    from rot2d3c_amd import closure as L
    B, H, W = 2, 8, 64
    rng = np.random.default_rng(43)
    net = _net(cuda, B, H, W, C.init_params(rng, scale=2.0))
    x = torch.from_numpy(rng.standard_normal((B, H, W, 3)).astype(np.float32)).to(cuda)
    y = torch.from_numpy(rng.standard_normal((B, H, W, 2)).astype(np.float32)).to(cuda)
    y0 = net.forward(x).cpu().numpy()
    l0, g0 = net.gradients(x, y)
    c = float(np.quantile(np.abs(y0), 0.7))
    net.set_output_clamp(c)
    y1 = net.forward(x).cpu().numpy()
    assert np.mean(np.abs(y0) > c) > 0.2
    assert np.array_equal(y1, np.clip(y0, -c, c))
    l1, g1 = net.gradients(x, y)
    assert l1 == l0 and torch.equal(g1, g0)
    net.set_output_clamp(0.0)
    assert np.array_equal(net.forward(x).cpu().numpy(), y0)
    # sign clipping: channel c zeroed where it has the sign of input channel c
    same = y0 * x.cpu().numpy()[..., :2] > 0
    assert 0.3 < same.mean() < 0.7
    for st in (1.0, 0.25):
        net.set_sign_clip(st)
        y2 = net.forward(x).cpu().numpy()
        keep = np.float32(1.0 - st)
        assert np.array_equal(y2, np.where(same, y0 * keep, y0).astype(np.float32))
    l2, g2 = net.gradients(x, y)
    assert l2 == l0 and torch.equal(g2, g0)
    net.set_sign_clip(0.0)
    with pytest.raises(Exception):
        net.set_sign_clip(1.5)
    with pytest.raises(Exception):
        net.set_output_clamp(-1.0)
    net.close()


@pytest.mark.parametrize("nsteps", [7, 20])
def test_ensemble_closure_period_tracks_the_stepwise_run(cuda, nsteps):
    """Closure forcing refreshed every 2nd step (held in between) inside the
    ensemble's step graphs against ClosureRun(period=2), and the held forcing
    changes the run against refreshing every step.  20 steps also run the
    steady steps as graphs of several steps."""
    print("ensemble closure period, %d steps" % nsteps, flush=True)
    from rot2d3c_amd import closure as L, ensemble as E
    n = 64
    ps0, vz0 = random_state(n, 37)
    from rot2d3c_amd import spectral as S
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    sin, sout = (0.02, 0.5, 3.0), (2.0, 0.2)
    spec = [dict(kf=6.0, f0=1.0, seed=9)]
    table = E.forcing_table(spec, nsteps)
    outs = {}
    for period in (2, 1):
        ps = torch.from_numpy(ps0).to(cuda)
        vz = torch.from_numpy(vz0).to(cuda)
        ens = E.Ensemble([E.Member(ps, vz, p, 1e-3)], schedules=None)
        net = L.Net(1, n, n, cuda, seed=8)
        ens.set_closure(0, net, sin, sout)
        ens.set_closure_period(period)
        ens.run(nsteps, table)
        torch.cuda.synchronize()
        outs[period] = (ps.cpu().numpy(), vz.cpu().numpy())
        ens.close()
    b = (torch.from_numpy(ps0).to(cuda), torch.from_numpy(vz0).to(cuda))
    run = L.ClosureRun(b[0], b[1], p, L.Net(1, n, n, cuda, seed=8), sin, sout, guard_every=0,
                       period=2)
    for st in range(nsteps):
        run.step(*table[st][0], dt=1e-3)
    torch.cuda.synchronize()
    assert rel_l2(outs[2][0], b[0].cpu().numpy()) < 1e-6
    assert rel_l2(outs[2][1], b[1].cpu().numpy()) < 1e-6
    assert rel_l2(outs[1][0], outs[2][0]) > 1e-9
