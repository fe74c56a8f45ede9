"""Numpy restatement of the closure network (test infrastructure only: the
checker of rot2d3c_amd/csrc/closure.hip, never called by the product).

The network: 6 periodic 3x3 convolutions, channels 3 -> 32 -> 32 -> 32 -> 32
-> 32 -> 2, GELU (erf form) after the first five, MSE loss over pixels and
both output channels.  Parameters are flat, layer by layer W[9][co][ci]
followed by b[co], with the input channels padded to 16 and the last layer's
output channels padded to 16 (as the HIP library stores them); tap
t = 3 (dy+1) + (dx+1) multiplies X[y+dy][x+dx].

`emulate=True` rounds to bfloat16 where the GPU does (weights, the input, each
layer's activation, each dZ) and accumulates in float64; `emulate=False` is the
exact float64 network, whose analytic gradient the tests check against
central finite differences (the pin of this restatement).
"""
import numpy as np
from scipy.special import erf

CI = [16, 32, 32, 32, 32, 32]
CO = [32, 32, 32, 32, 32, 16]
NIN, NOUT = 3, 2


def offsets():
    """(weight offset, bias offset) of each layer in the flat parameters."""
    out, o = [], 0
    for ci, co in zip(CI, CO):
        out.append((o, o + 9 * co * ci))
        o += 9 * co * ci + co
    return out, o


def unpack(params):
    offs, _ = offsets()
    Ws, bs = [], []
    for (wo, bo), ci, co in zip(offs, CI, CO):
        Ws.append(params[wo:wo + 9 * co * ci].reshape(9, co, ci))
        bs.append(params[bo:bo + co])
    return Ws, bs


def init_params(rng, scale=1.0):
    """He-style random parameters with the padding zero (float32 values)."""
    offs, total = offsets()
    p = np.zeros(total, dtype=np.float64)
    for l, ((wo, bo), ci, co) in enumerate(zip(offs, CI, CO)):
        cin = NIN if l == 0 else ci
        cout = NOUT if l == len(CI) - 1 else co
        w = np.zeros((9, co, ci))
        w[:, :cout, :cin] = rng.standard_normal((9, cout, cin)) * scale * np.sqrt(2.0 / (9 * cin))
        p[wo:wo + 9 * co * ci] = w.ravel()
        p[bo:bo + cout] = 0.01 * scale * rng.standard_normal(cout)
    return p.astype(np.float32).astype(np.float64)


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, as the kernels do."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _taps():
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(X, W, b):
    """Z[b,y,x,co] = sum_t sum_ci W[t,co,ci] X[b, y+dy, x+dx, ci] + b[co]."""
    Z = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, (dy, dx) in enumerate(_taps()):
        Z += np.roll(X, (-dy, -dx), axis=(1, 2)) @ W[t].T
    return Z + (0.0 if b is None else b)


def forward(params, inp, emulate=True):
    """inp [B,H,W,3] -> (prediction [B,H,W,2], cache)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unpack(params)
    X = np.zeros(inp.shape[:3] + (16,))
    X[..., :NIN] = inp
    X = r(X)
    Xs, Zs = [X], []
    for l in range(6):
        Z = conv(X, r(Ws[l]), bs[l])
        if emulate:
            Z = Z.astype(np.float32).astype(np.float64)
        Zs.append(Z)
        X = r(gelu(Z)) if l < 5 else r(Z)
        Xs.append(X)
    return Zs[5][..., :NOUT], (Xs, Zs)


def loss_and_grad(params, inp, tgt, emulate=True):
    """MSE loss and its gradient with respect to the flat parameters."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, _ = unpack(params)
    pred, (Xs, Zs) = forward(params, inp, emulate)
    P = pred.shape[0] * pred.shape[1] * pred.shape[2]
    d = pred - tgt
    if emulate:
        d = (pred.astype(np.float32) - tgt.astype(np.float32)).astype(np.float64)
    loss = float(np.sum(d * d) / (2.0 * P))
    G = np.zeros(Zs[5].shape)
    G[..., :NOUT] = d * (1.0 / P)
    G = r(G)
    offs, total = offsets()
    grad = np.zeros(total)
    for l in range(5, -1, -1):
        wo, bo = offs[l]
        X = Xs[l]
        dW = np.zeros((9, CO[l], CI[l]))
        for t, (dy, dx) in enumerate(_taps()):
            Xs_t = np.roll(X, (-dy, -dx), axis=(1, 2))
            dW[t] = np.einsum("bhwo,bhwi->oi", G, Xs_t)
        grad[wo:bo] = dW.ravel()
        grad[bo:bo + CO[l]] = G.sum(axis=(0, 1, 2))
        if l == 0:
            break
        Wr = r(Ws[l])
        dX = np.zeros(X.shape)
        for t, (dy, dx) in enumerate(_taps()):
            dX += np.roll(G @ Wr[t], (dy, dx), axis=(1, 2))
        G = r(dX * gelu_grad(Zs[l - 1]))
    return loss, grad
