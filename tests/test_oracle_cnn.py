"""Pins of the closure network's numpy restatement (oracle/cnn.py): the
convolution against a brute-force sum, the float64 backprop against central
finite differences, GELU' against a numerical derivative, and bf16 rounding
against known values."""
import numpy as np
import pytest

from oracle import cnn as C


def test_conv_equals_brute_force_periodic_sum():
    rng = np.random.default_rng(0)
    B, H, W, ci, co = 2, 4, 5, 3, 2
    X = rng.standard_normal((B, H, W, ci))
    Wt = rng.standard_normal((9, co, ci))
    b = rng.standard_normal(co)
    Z = C.conv(X, Wt, b)
    ref = np.zeros((B, H, W, co))
    for bb in range(B):
        for y in range(H):
            for x in range(W):
                for o in range(co):
                    s = b[o]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            t = 3 * (dy + 1) + (dx + 1)
                            s += Wt[t, o] @ X[bb, (y + dy) % H, (x + dx) % W]
                    ref[bb, y, x, o] = s
    assert np.allclose(Z, ref, rtol=1e-13, atol=1e-13)


def test_gelu_derivative():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    num = (C.gelu(z + h) - C.gelu(z - h)) / (2 * h)
    assert np.allclose(C.gelu_grad(z), num, atol=1e-8)
    assert C.gelu(np.array(0.0)) == 0.0 and abs(C.gelu(np.array(10.0)) - 10.0) < 1e-12


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e-3])
    got = C.bf16(v)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6
    assert list(got[:4]) == [1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7]
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


def test_float64_gradient_equals_finite_differences():
    """The analytic backprop of the exact (float64) network against central
    differences of its loss, on random parameters, inputs and targets; the
    padding's gradient is exactly zero."""
    rng = np.random.default_rng(1)
    p = C.init_params(rng, scale=1.0)
    inp = rng.standard_normal((2, 4, 6, 3))
    tgt = rng.standard_normal((2, 4, 6, 2))
    loss, g = C.loss_and_grad(p, inp, tgt, emulate=False)
    offs, total = C.offsets()
    live = np.flatnonzero(p != 0)
    idx = rng.choice(live, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(total)
        e[i] = h
        num = (C.loss_and_grad(p + e, inp, tgt, emulate=False)[0]
               - C.loss_and_grad(p - e, inp, tgt, emulate=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, num, g[i])
    # padded input channels of layer 0 and padded outputs of layer 5
    W0 = g[offs[0][0]:offs[0][1]].reshape(9, 32, 16)
    assert np.all(W0[:, :, 3:] == 0)
    W5 = g[offs[5][0]:offs[5][1]].reshape(9, 16, 32)
    assert np.all(W5[:, 2:, :] == 0) and np.all(g[offs[5][1] + 2:offs[5][1] + 16] == 0)


def test_emulated_network_stays_close_to_exact():
    """bf16 rounding perturbs the prediction by O(2^-8) relative, no more."""
    rng = np.random.default_rng(2)
    p = C.init_params(rng)
    inp = rng.standard_normal((1, 4, 8, 3))
    a, _ = C.forward(p, inp, emulate=True)
    b, _ = C.forward(p, inp, emulate=False)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-2
