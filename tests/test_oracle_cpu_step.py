"""The compiled CPU restatement of the step (oracle/cpu_step.cpp, the
benchmark's CPU baseline) against the numpy oracle: its own 2D transforms
to 1e-13 and forced fixed-dt steps to the FP64 bar of 1e-12 relative L2."""
import numpy as np
import pytest

from conftest import random_state, rel_l2
from oracle import cpu_step as C
from oracle import spectral as O


@pytest.fixture(scope="module", autouse=True)
def built():
    C.build()
    C.threads(4)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 512])
def test_transforms_match_the_oracle(n):
    rng = np.random.default_rng(n)
    spec = (rng.standard_normal((n // 2 + 1, n)) + 1j * rng.standard_normal((n // 2 + 1, n))) * n
    assert rel_l2(C.c2r(spec), O.fftp2d_complex_to_real(spec)) < 1e-13
    real = rng.standard_normal((n, n))
    assert rel_l2(C.r2c(real), O.fftp2d_real_to_complex(real)) < 1e-13


def _draws(kup, seed, steps):
    s0, s1 = seed, seed + 1
    _, _, _, s0 = O.rand_force_draw(kup, s0)  # the pre-loop pair (src/Rot2D3C.para.f90:377-380)
    _, _, _, s1 = O.rand_force_draw(kup, s1)
    out = []
    for _ in range(steps):
        kx, ky, ph, s0 = O.rand_force_draw(kup, s0)
        a = (kx, ky, ph, 1.0, kup)
        kx, ky, ph, s1 = O.rand_force_draw(kup, s1)
        out.append((a, (kx, ky, ph, 1.0, 1.0)))
    return out


@pytest.mark.parametrize("n", [16, 32, 64, 256])
def test_forced_steps_match_the_oracle(n):
    ps, vz = random_state(n, 3)
    p = O.Params(1e-3, 0.1, 1, 2, 2e-3, 0.05, 2, 1, 2.0)
    steps, dt, kup, seed = 3, 2e-3, n / 8.0, 7
    ops, ovz, *_ = O.run_random_forcing(ps.copy(), vz.copy(), steps, 0.0, p, kup, kup, 1.0, 1.0,
                                        seed, fixed_dt=dt)
    cps, cvz = ps.copy(), vz.copy()
    C.run(cps, cvz, p, steps, dt, _draws(kup, seed, steps))
    assert rel_l2(cps, ops) < 1e-12
    assert rel_l2(cvz, ovz) < 1e-12
