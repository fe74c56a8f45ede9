"""MI355X-native Rot2D3C pseudospectral right-hand side and RK integrator.

`rot2d3c_amd.spectral` holds the reference-named operators (derivk2, poisson,
fftp2d_complex_to_real, CFL_condition, rk4_step, ...) on device tensors; they
run the HIP kernels of librot2d3c.so through the C ABI of include/rot2d3c.h.
"""
from ._lib import LIB_PATH, R2D3CError  # noqa: F401

__version__ = "0.1.0"
