import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def random_state(n, seed, e_psi=1.0, e_vz=1.0):
    """Band-limited random psi, vz (dealiased, Hermitian by construction),
    normalised to energy(ps,1) = e_psi and energy(vz,0) = e_vz."""
    from oracle import spectral as O
    rng = np.random.default_rng(seed)
    g = O.grid(n)
    ps = O.fftp2d_real_to_complex(rng.standard_normal((n, n)))
    vz = O.fftp2d_real_to_complex(rng.standard_normal((n, n)))
    ps[~g.dealias] = 0
    vz[~g.dealias] = 0
    ps[0, 0] = 0
    vz[0, 0] = 0
    ps *= np.sqrt(e_psi / O.energy(ps, 1))
    vz *= np.sqrt(e_vz / O.energy(vz, 0))
    return ps, vz


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
