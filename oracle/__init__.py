"""CPU oracle for the Rot2D3C pseudospectral right-hand side and RK integrator.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference Fortran (src/Rot2D3C.para.f90, src/pseudospec2D_Rot2D3C.f90,
src/pseudospec2D_mod.f90, src/fftp2D.f90), plus a compiled C++/OpenMP
restatement of the time step (cpu_step.cpp, wrapped by cpu_step.py; checked
against the numpy one by tests/test_oracle_cpu_step.py), which is the
benchmark's CPU baseline.  It is the checker: only `tests/`,
`__graft_entry__.smoke()` (and `build()`, which compiles cpu_step.cpp) and the
`cpu_baseline` leg of `bench.py` may import it.  The product (`rot2d3c_amd`)
never imports, calls or falls back to it.

Pinning status of each function (what checks it independently of its own
formula) is recorded in the module docstring of `oracle.spectral`.
"""
from . import spectral  # noqa: F401
