"""Learned subgrid closure for a coarse grid (no reference counterpart).

A coarse run (e.g. 256^2) stands in for a fine DNS (e.g. 2048^2).  The
closure predicts, from the coarse fields, the part of the nonlinear terms the
coarse grid misses: with the coarse-graining G (box filter of width one coarse
cell, then truncation to the coarse grid's dealiased modes),

    tau_5 = G(poisson(psi, laplak2(psi))) - poisson(G psi, laplak2(G psi))
    tau_6 = G(poisson(psi, vz))           - poisson(G psi, G vz)

the subgrid parts of C5 and C6 of the stage update (src/Rot2D3C.para.f90:
571-576,585-592).  The linear terms (rotation coupling, dissipation, forcing)
commute with G and need no closure.  In the coarse run the closure enters the
update like a forcing: fp += -tau_5/k^2, fz += tau_6 (r2d3c_closure_forcing).
It is evaluated once per time step from the step's initial state and held
for the four stages.

The network (r2d3c_cnn: 6 periodic 3x3 convolutions with 32 channels and GELU,
BF16 MFMA with FP32 accumulation) maps the real-space (laplacian of psi, vz,
psi) to real-space (tau_5, tau_6), each channel normalised by its standard
deviation over the training set.  It is trained on the GPU by the library's
own forward and backward kernels with Adam.  The plain coarse solver is the
fallback: ClosureRun switches to it when the closure run stops being finite
or its energy leaves a bound.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import spectral as S
from ._lib import check


def _lib_():
    return _lib.lib()


def _fptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _dvec(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


# --- the network ----------------------------------------------------------
def layer_layout():
    n = 6
    wo, bo = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    ci, co = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    _lib_().r2d3c_cnn_layer_offsets(wo, bo, ci, co)
    return list(wo), list(bo), list(ci), list(co)


def init_params(seed=0, scale=1.0):
    """He-normal weights on the live channels (3 inputs, 2 outputs), small
    biases, zero padding; float32 numpy array in the library's layout."""
    rng = np.random.default_rng(seed)
    wo, bo, ci, co = layer_layout()
    p = np.zeros(_lib_().r2d3c_cnn_param_count(), dtype=np.float32)
    for l in range(6):
        cin = 3 if l == 0 else ci[l]
        cout = 2 if l == 5 else co[l]
        w = np.zeros((9, co[l], ci[l]), dtype=np.float32)
        w[:, :cout, :cin] = rng.standard_normal((9, cout, cin)) * scale * math.sqrt(2.0 / (9 * cin))
        p[wo[l]:wo[l] + w.size] = w.ravel()
        p[bo[l]:bo[l] + cout] = 0.01 * scale * rng.standard_normal(cout)
    return p


class Net:
    """The closure CNN for batches of up to `batch` images of h x w."""

    def __init__(self, batch, height, width, device, params=None, seed=0):
        self.batch, self.h, self.w = int(batch), int(height), int(width)
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib_().r2d3c_cnn_create(ctypes.byref(h), self.batch, self.h, self.w),
                  "r2d3c_cnn_create")
        self.handle = h
        self.loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.set_params(init_params(seed) if params is None else params)

    def _stream(self):
        return S._stream(self.device)

    def set_params(self, params):
        t = torch.as_tensor(np.asarray(params, dtype=np.float32)).to(self.device)
        check(_lib_().r2d3c_cnn_set_params(self.handle, _fptr(t), self._stream()),
              "r2d3c_cnn_set_params")
        torch.cuda.synchronize(self.device)

    def params(self):
        t = torch.empty(_lib_().r2d3c_cnn_param_count(), dtype=torch.float32, device=self.device)
        check(_lib_().r2d3c_cnn_get_params(self.handle, _fptr(t), self._stream()),
              "r2d3c_cnn_get_params")
        return t

    def _check(self, x, ch):
        if (x.dtype != torch.float32 or not x.is_contiguous() or x.device != self.device
                or x.dim() != 4 or x.shape[1:] != (self.h, self.w, ch) or x.shape[0] > self.batch):
            raise ValueError("expected float32 [B<=%d, %d, %d, %d] on %s"
                             % (self.batch, self.h, self.w, ch, self.device))

    def forward(self, x, out=None):
        self._check(x, 3)
        out = torch.empty(x.shape[:3] + (2,), dtype=torch.float32, device=self.device) \
            if out is None else out
        check(_lib_().r2d3c_cnn_forward(self.handle, _fptr(x), x.shape[0], _fptr(out),
                                        self._stream()), "r2d3c_cnn_forward")
        return out

    def set_forcing_gain(self, gain=None):
        """Gain per shell int(|k| + 0.5) of the closure forcing made from this
        network (array indexed by shell; None: no gain).  Set it before the
        network is attached to an ensemble run."""
        if gain is None:
            check(_lib_().r2d3c_cnn_set_forcing_gain(self.handle, None, 0),
                  "r2d3c_cnn_set_forcing_gain")
            return
        g = np.asarray(gain, dtype=np.float64)
        check(_lib_().r2d3c_cnn_set_forcing_gain(self.handle, _dvec(g), len(g)),
              "r2d3c_cnn_set_forcing_gain")

    def set_output_clamp(self, clamp):
        """Saturate inference outputs to [-clamp, clamp] (0: none); training
        and gradients are unaffected."""
        check(_lib_().r2d3c_cnn_set_output_clamp(self.handle, float(clamp)),
              "r2d3c_cnn_set_output_clamp")

    def set_sign_clip(self, strength):
        """Inference only: scale output channel c (c = 0, 1) by 1 - strength
        wherever it has the sign of input channel c (strength 1: the
        prediction only damps those fields pointwise; 0: off)."""
        check(_lib_().r2d3c_cnn_set_sign_clip(self.handle, float(strength)),
              "r2d3c_cnn_set_sign_clip")

    def gradients(self, x, y):
        self._check(x, 3)
        self._check(y, 2)
        g = torch.empty(_lib_().r2d3c_cnn_param_count(), dtype=torch.float32, device=self.device)
        check(_lib_().r2d3c_cnn_gradients(self.handle, _fptr(x), _fptr(y), x.shape[0], _fptr(g),
                                          _fptr(self.loss), self._stream()), "r2d3c_cnn_gradients")
        return float(self.loss.item()), g

    def train_step(self, x, y, lr):
        """One Adam step; the loss stays on the device (self.loss)."""
        check(_lib_().r2d3c_cnn_train_step(self.handle, _fptr(x), _fptr(y), x.shape[0],
                                           ctypes.c_float(lr), _fptr(self.loss), self._stream()),
              "r2d3c_cnn_train_step")
        return self.loss

    def close(self):
        if self.handle:
            _lib_().r2d3c_cnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- coarse-graining and training data -------------------------------------
def coarse_grain(a_fine, n_coarse, width=None):
    """G(a): box filter of width one coarse cell (2 pi / n_coarse unless given)
    and truncation to the coarse grid's dealiased modes."""
    nf = S._spec_n(a_fine)
    width = 2.0 * math.pi / n_coarse if width is None else width
    pf, pc = S.Plan.get(nf, a_fine.device), S.Plan.get(n_coarse, a_fine.device)
    out = torch.empty((n_coarse // 2 + 1, n_coarse), dtype=torch.complex128, device=a_fine.device)
    check(_lib_().r2d3c_filter(pf.handle, pc.handle, S._ptr(a_fine), S._ptr(out), float(width),
                               S._stream(a_fine.device)), "r2d3c_filter")
    return out


def box_gain(kx, ky, n_coarse, width=None):
    """The coarse-graining's transfer function at wavevector (kx, ky):
    sinc(kx w/2) sinc(ky w/2), w the filter width (one coarse cell)."""
    w = 2.0 * math.pi / n_coarse if width is None else width
    g = 1.0
    for k in (kx, ky):
        a = 0.5 * float(k) * w
        g *= 1.0 if a == 0.0 else math.sin(a) / a
    return g


def filtered_draw(draw, n_coarse, width=None):
    """The rand_force draw `draw` (spectral.ForceDraw) coarse-grained: the
    filtered equation is forced by G f, whose amplitude is G times the draw's,
    i.e. injection rate f0 G^2 (the wavevector and phase are unchanged)."""
    if draw is None:
        return None
    g = box_gain(draw.kx, draw.ky, n_coarse, width)
    return S.ForceDraw(draw.kx, draw.ky, draw.phase, draw.f0 * g * g, draw.kupk)


def network_inputs(ps, vz, scale=(1.0, 1.0, 1.0), out=None):
    """fp32 [n, n, 3]: real-space (laplacian of psi, vz, psi) times scale."""
    n = S._spec_n(ps)
    out = torch.empty((n, n, 3), dtype=torch.float32, device=ps.device) if out is None else out
    pl = S.Plan.get(n, ps.device)
    check(_lib_().r2d3c_closure_inputs(pl.handle, S._ptr(ps), S._ptr(vz), _dvec(scale), _fptr(out),
                                       S._stream(ps.device)), "r2d3c_closure_inputs")
    return out


def snapshot(ps_f, vz_f, n_coarse, width=None, kmin=0.0):
    """(coarse psi, coarse vz, inputs [n,n,3], targets [n,n,2]) of one fine
    state; the targets keep the modes |k| >= kmin."""
    ps_c = coarse_grain(ps_f, n_coarse, width)
    vz_c = coarse_grain(vz_f, n_coarse, width)
    f5 = coarse_grain(S.poisson(ps_f, S.laplak2(ps_f)), n_coarse, width)
    f6 = coarse_grain(S.poisson(ps_f, vz_f), n_coarse, width)
    c5 = S.poisson(ps_c, S.laplak2(ps_c))
    c6 = S.poisson(ps_c, vz_c)
    tgt = torch.empty((n_coarse, n_coarse, 2), dtype=torch.float32, device=ps_f.device)
    pl = S.Plan.get(n_coarse, ps_f.device)
    check(_lib_().r2d3c_closure_targets(pl.handle, S._ptr(f5), S._ptr(c5), S._ptr(f6), S._ptr(c6),
                                        _dvec((1.0, 1.0)), float(kmin), _fptr(tgt),
                                        S._stream(ps_f.device)),
          "r2d3c_closure_targets")
    return ps_c, vz_c, network_inputs(ps_c, vz_c), tgt


def closure_forcing(pred, n, scale, fp=None, fz=None, kmin=0.0):
    """Dense (fp, fz) of a prediction [n, n, 2] times scale (the de-normalisation),
    on the modes |k| >= kmin."""
    dev = pred.device
    fp = torch.empty((n // 2 + 1, n), dtype=torch.complex128, device=dev) if fp is None else fp
    fz = torch.empty_like(fp) if fz is None else fz
    pl = S.Plan.get(n, dev)
    check(_lib_().r2d3c_closure_forcing(pl.handle, _fptr(pred), _dvec(scale), float(kmin),
                                        S._ptr(fp), S._ptr(fz), S._stream(dev)),
          "r2d3c_closure_forcing")
    return fp, fz


def clip_forcing(ps, vz, fp, fz, kclip=1.0):
    """Backscatter clipping of a closure forcing (fp, fz), in place, on the
    modes |k| >= kclip (default: all): where it injects energy
    (Re(conj(psi) fp) > 0, Re(conj(vz) fz) > 0) its component along the
    mode's coefficient is removed."""
    n = S._spec_n(ps)
    for t in (vz, fp, fz):
        if S._spec_n(t) != n:
            raise ValueError("clip_forcing: arrays of one grid size expected")
    pl = S.Plan.get(n, ps.device)
    check(_lib_().r2d3c_closure_clip(pl.handle, S._ptr(ps), S._ptr(vz), float(kclip), S._ptr(fp),
                                     S._ptr(fz), S._stream(ps.device)), "r2d3c_closure_clip")
    return fp, fz


# --- training ----------------------------------------------------------------
def taper_gain(kmin, ktaper, nshells):
    """Closure forcing gain rising linearly from 0 at shell kmin to 1 at
    shell ktaper (a hard cutoff at kmin when ktaper <= kmin)."""
    k = np.arange(nshells, dtype=np.float64)
    if ktaper <= kmin:
        return np.where(k >= kmin, 1.0, 0.0)
    return np.clip((k - kmin) / (ktaper - kmin), 0.0, 1.0)


def channel_std(x):
    """Per-channel standard deviation of [N, h, w, C] (float64 host values)."""
    flat = x.reshape(-1, x.shape[-1]).double()
    return [float(v) for v in flat.std(dim=0).cpu()]


def train(net, X, Y, epochs, lr=1e-3, lr_end=1e-4, seed=0, log=None):
    """Adam over shuffled minibatches of the normalised data X [N,h,w,3],
    Y [N,h,w,2] (device float32); cosine learning-rate decay.  Returns the mean
    training loss of each epoch."""
    rng = np.random.default_rng(seed)
    N, B = X.shape[0], net.batch
    steps = epochs * (N // B)
    hist, k = [], 0
    xb = torch.empty((B,) + tuple(X.shape[1:]), dtype=torch.float32, device=X.device)
    yb = torch.empty((B,) + tuple(Y.shape[1:]), dtype=torch.float32, device=Y.device)
    for ep in range(epochs):
        perm = torch.as_tensor(rng.permutation(N), device=X.device)
        acc = torch.zeros(1, dtype=torch.float64, device=X.device)
        for i in range(N // B):
            idx = perm[i * B:(i + 1) * B]
            torch.index_select(X, 0, idx, out=xb)
            torch.index_select(Y, 0, idx, out=yb)
            a = k / max(1, steps - 1)
            rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * a))
            acc += net.train_step(xb, yb, rate)
            k += 1
        hist.append(float(acc.item()) / (N // B))
        if log:
            log("epoch %d loss %.5f" % (ep, hist[-1]))
    return hist


def evaluate(net, X, Y):
    """Mean loss of the network on X, Y (no update)."""
    B, tot, cnt = net.batch, 0.0, 0
    for i in range(0, X.shape[0] - B + 1, B):
        loss, _ = net.gradients(X[i:i + B].contiguous(), Y[i:i + B].contiguous())
        tot += loss
        cnt += 1
    return tot / max(cnt, 1)


# --- run-time stabiliser ------------------------------------------------------
def cfl_number(ps, vz, p, dt):
    """CFL number of a fixed step dt: dt over the CFL condition's step at CFL
    number 1 (the reference's formula; NaN for a non-finite state)."""
    d1 = S.CFL_condition(1.0, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
    return dt / d1 if math.isfinite(d1) and d1 > 0.0 else float("nan")


class SafeMode:
    """Run-time stabiliser of closure runs with a fixed step.

    A learned closure can feed a compact vortex until the fixed step's CFL
    number grows without bound, although energy and enstrophy stay normal.
    Checked every few steps (check()), a closure run whose CFL number exceeds
    cfl_on continues with its network fully sign-clipped (it can only damp
    vorticity and vz pointwise, Net.set_sign_clip(1)) until the CFL number
    falls below cfl_off, then returns to the network's own setting.

    apply(net, flag) sets a network's mode; for a run inside an ensemble the
    caller re-registers the network (Ensemble.set_closure) so the step graph
    picks the change up."""

    def __init__(self, cfl_on, cfl_off, base_sign_clip=0.0):
        if not cfl_off <= cfl_on:
            raise ValueError("SafeMode: cfl_off must not exceed cfl_on")
        self.cfl_on, self.cfl_off, self.base = float(cfl_on), float(cfl_off), float(base_sign_clip)
        self.active = {}
        self.checks = 0
        self.active_checks = 0

    def apply(self, net, flag):
        net.set_sign_clip(1.0 if flag else self.base)

    def check(self, run, cfl):
        """Record run's CFL number; returns True when its mode must change
        (then self.active[run] is the new mode)."""
        was = self.active.get(run, False)
        self.checks += 1
        self.active_checks += was
        now = cfl > self.cfl_on if not was else not cfl < self.cfl_off
        self.active[run] = now
        return now != was

    @property
    def fraction(self):
        """Fraction of checks at which a run was in the damping mode."""
        return self.active_checks / max(self.checks, 1)


# --- the coarse run with the closure -----------------------------------------
class ClosureRun:
    """Coarse run of psi, vz (advanced in place) with the learned closure.

    scale_in / scale_out: the per-channel normalisation of the network's
    inputs and outputs (1/std of the training inputs, std of the targets).
    With net=None (or after a fallback) it is the plain coarse solver; with
    kclip > 0 the closure forcing is backscatter-clipped on |k| >= kclip
    (clip_forcing); the forcing is evaluated on every period-th step (counted
    from construction) and held in between."""

    def __init__(self, ps, vz, params, net=None, scale_in=None, scale_out=None, guard_every=50,
                 energy_bound=None, kmin=0.0, kclip=0.0, period=1):
        self.ps, self.vz, self.p = ps, vz, params
        self.n = S._spec_n(ps)
        self.net = net
        self.scale_in, self.scale_out = scale_in, scale_out
        self.kmin, self.kclip = kmin, float(kclip)
        self.period = int(period)  # forcing evaluated every period-th step, held between
        self.active = net is not None
        self.guard_every = guard_every
        self.energy_bound = energy_bound
        self.fallbacks = 0
        self._steps = 0
        dev = ps.device
        self._inp = torch.empty((1, self.n, self.n, 3), dtype=torch.float32, device=dev)
        self._pred = torch.empty((1, self.n, self.n, 2), dtype=torch.float32, device=dev)
        self._fp = torch.empty_like(ps)
        self._fz = torch.empty_like(ps)
        self._good = (ps.clone(), vz.clone())

    def step(self, fp_draw=None, fz_draw=None, dt=None, cfl=None, dt_out=None):
        dense = None
        if self.active and self._steps % self.period:
            dense = (self._fp, self._fz)  # the held forcing
        elif self.active:
            network_inputs(self.ps, self.vz, self.scale_in, out=self._inp[0])
            self.net.forward(self._inp, out=self._pred)
            closure_forcing(self._pred[0], self.n, self.scale_out, self._fp, self._fz, self.kmin)
            if self.kclip > 0.0:
                clip_forcing(self.ps, self.vz, self._fp, self._fz, self.kclip)
            dense = (self._fp, self._fz)
        S.rk4_step(self.ps, self.vz, self.p, fp_draw, fz_draw, dt=dt, cfl=cfl, dt_out=dt_out,
                   dense=dense)
        self._steps += 1
        if self.active and self.guard_every and self._steps % self.guard_every == 0:
            self._guard()

    def _guard(self):
        e = S.energy(self.ps, 1) + S.energy(self.vz, 0)
        bad = not math.isfinite(e) or (self.energy_bound is not None and e > self.energy_bound)
        if bad:
            # fall back to the plain coarse solver from the last good state
            self.ps.copy_(self._good[0])
            self.vz.copy_(self._good[1])
            self.active = False
            self.fallbacks += 1
        else:
            self._good[0].copy_(self.ps)
            self._good[1].copy_(self.vz)
