"""r2d3c_run_noalloc (no device allocation, static scratch array) against
r2d3c_run on a plan, through the C ABI: two calls in a row on one device, the
second at a smaller grid, each bit for bit equal to the plan path (the same
kernels on the same data), dt included."""
import ctypes

import numpy as np
import pytest

from conftest import random_state
from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _draws(kup, seed, steps):
    from rot2d3c_amd import spectral as S
    out = []
    s0, s1 = seed, seed + 1
    for _ in range(steps):
        kx, ky, ph, s0 = S.rand_force_draw(kup, s0)
        a = S.ForceDraw(kx, ky, ph, 1.0, kup)
        kx, ky, ph, s1 = S.rand_force_draw(kup, s1)
        b = S.ForceDraw(kx, ky, ph, 1.0, 1.0)
        out.append((a, b))
    return out


def _run_both(n, steps, cfl_mode, seed, cuda):
    import torch
    from rot2d3c_amd import _lib
    from rot2d3c_amd import spectral as S
    print("noalloc n=%d steps=%d cfl=%d seed=%d" % (n, steps, cfl_mode, seed), flush=True)
    ps, vz = random_state(n, seed)
    p = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    draws = _draws(n / 32.0 + 2, seed, steps)
    L = _lib.lib()
    arr = (_lib.Forcing * steps)()
    for k, (fa, fb) in enumerate(draws):
        arr[k] = S._forcing_c(fa, fb)
    cp = p.c()
    mode, val = (_lib.R2D3C_DT_CFL, 0.5) if cfl_mode else (_lib.R2D3C_DT_FIXED, 2e-3)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    # the existing entry point on a new plan (default schedules, as the
    # no-alloc entry point uses)
    a_ps = torch.from_numpy(ps).to(cuda)
    a_vz = torch.from_numpy(vz).to(cuda)
    a_dt = torch.zeros(steps, dtype=torch.float64, device=cuda)
    plan = ctypes.c_void_p()
    assert L.r2d3c_plan_create(ctypes.byref(plan), n) == 0, L.r2d3c_last_error()
    try:
        rc = L.r2d3c_run(plan, ctypes.byref(cp), ctypes.c_void_p(a_ps.data_ptr()),
                         ctypes.c_void_p(a_vz.data_ptr()), arr, steps, mode, float(val),
                         ctypes.c_void_p(a_dt.data_ptr()), ctypes.c_void_p(stream))
        assert rc == 0, L.r2d3c_last_error()
        torch.cuda.synchronize(cuda)
    finally:
        L.r2d3c_plan_destroy(plan)
    b_ps = torch.from_numpy(ps).to(cuda)
    b_vz = torch.from_numpy(vz).to(cuda)
    b_dt = torch.zeros(steps, dtype=torch.float64, device=cuda)
    rc = _lib.lib().r2d3c_run_noalloc(n, ctypes.byref(cp), ctypes.c_void_p(b_ps.data_ptr()),
                                      ctypes.c_void_p(b_vz.data_ptr()), arr, steps, mode,
                                      float(val), ctypes.c_void_p(b_dt.data_ptr()),
                                      ctypes.c_void_p(stream))
    assert rc == 0, _lib.lib().r2d3c_last_error()
    torch.cuda.synchronize(cuda)
    return (a_ps.cpu().numpy(), a_vz.cpu().numpy(), a_dt.cpu().numpy(),
            b_ps.cpu().numpy(), b_vz.cpu().numpy(), b_dt.cpu().numpy())


def test_noalloc_equals_plan_run_twice_smaller_second(cuda):
    for n, steps, cfl_mode, seed in ((512, 4, 1, 11), (128, 3, 0, 12), (1024, 2, 1, 13),
                                     (256, 5, 1, 14)):
        a_ps, a_vz, a_dt, b_ps, b_vz, b_dt = _run_both(n, steps, cfl_mode, seed, cuda)
        assert np.all(np.isfinite(b_ps)) and np.all(np.isfinite(b_vz))
        assert np.array_equal(a_dt, b_dt), (n, a_dt, b_dt)
        assert np.array_equal(a_ps, b_ps), n
        assert np.array_equal(a_vz, b_vz), n


def test_noalloc_matches_oracle(cuda):
    """One forced CFL step at 64 against the oracle (FP64 bar 1e-12)."""
    from conftest import rel_l2
    a_ps, a_vz, a_dt, b_ps, b_vz, b_dt = _run_both(64, 1, 1, 21, cuda)
    ps, vz = random_state(64, 21)
    p = O.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0)
    dt = O.CFL_condition(0.5, ps, vz, p.nu, p.nn, p.nuv, p.nnv, p.omega)
    assert abs(b_dt[0] - dt) <= 1e-13 * dt
    kup = 64 / 32.0 + 2
    fp, _ = O.rand_force(kup, kup, 1.0, b_dt[0], 21, 1, 64)
    fz, _ = O.rand_force(kup, kup, 1.0, b_dt[0], 22, 0, 64)
    eps, evz = O.rk_stages(ps, vz, fp, fz, b_dt[0], p)
    assert rel_l2(b_ps, eps) < 1e-12
    assert rel_l2(b_vz, evz) < 1e-12


def test_noalloc_rejects_sizes_beyond_the_maximum(cuda):
    import torch
    from rot2d3c_amd import _lib
    from rot2d3c_amd import spectral as S
    n = 2048
    t = torch.zeros((n // 2 + 1, n), dtype=torch.complex128, device=cuda)
    u = torch.zeros_like(t)
    cp = S.Params(1e-3, 0.1, 1, 2, 1e-3, 0.1, 1, 2, 2.0).c()
    rc = _lib.lib().r2d3c_run_noalloc(n, ctypes.byref(cp), ctypes.c_void_p(t.data_ptr()),
                                      ctypes.c_void_p(u.data_ptr()), None, 1,
                                      _lib.R2D3C_DT_FIXED, 1e-3, None, None)
    assert rc == -1
    assert b"power of two in [16, 1024]" in _lib.lib().r2d3c_last_error()
