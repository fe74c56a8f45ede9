"""ctypes wrapper of oracle/cpu_step.cpp, the compiled CPU restatement of the
Rot2D3C time step (test infrastructure: the benchmark's CPU baseline; the
product never loads it).  build() compiles it with g++ and OpenMP into
oracle/libcpu_step.so."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_step.cpp")
LIB = os.path.join(HERE, "libcpu_step.so")
CMD = ["g++", "-O3", "-std=c++17", "-fcx-fortran-rules", "-fopenmp", "-fPIC", "-shared", "-o", LIB, SRC]

_lib = None


def build(verbose=False):
    """Compile libcpu_step.so (always recompiles)."""
    if verbose:
        print(" ".join(CMD), flush=True)
    subprocess.run(CMD, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/libcpu_step.so is missing: run oracle.cpu_step.build()")
        _lib = ctypes.CDLL(LIB)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        _lib.r2d3c_cpu_threads.argtypes, _lib.r2d3c_cpu_threads.restype = [i], i
        _lib.r2d3c_cpu_c2r.argtypes, _lib.r2d3c_cpu_c2r.restype = [i, p, p], i
        _lib.r2d3c_cpu_r2c.argtypes, _lib.r2d3c_cpu_r2c.restype = [i, p, p], i
        _lib.r2d3c_cpu_run.argtypes = [i, p, p, p, i, d, p]
        _lib.r2d3c_cpu_run.restype = i
    return _lib


def threads(n=0):
    """Set (n > 0) and return the OpenMP thread count of the restatement."""
    return int(lib().r2d3c_cpu_threads(int(n)))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def c2r(spec):
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    n = spec.shape[1]
    out = np.empty((n, n))
    lib().r2d3c_cpu_c2r(n, _ptr(spec), _ptr(out))
    return out


def r2c(real):
    real = np.ascontiguousarray(real, dtype=np.float64)
    n = real.shape[0]
    out = np.empty((n // 2 + 1, n), dtype=np.complex128)
    lib().r2d3c_cpu_r2c(n, _ptr(real), _ptr(out))
    return out


def run(ps, vz, p, steps, dt, draws=None):
    """`steps` fixed-dt steps in place on complex128 arrays ps, vz.  p: an
    oracle Params; draws: per step a pair of (kx, ky, phase, f0, kupk) tuples
    for psi and vz (None entries: no forcing), or None."""
    n = ps.shape[1]
    prm = np.array([p.nu, p.hnu, p.nuv, p.hnuv, p.omega, p.nn, p.mm, p.nnv, p.mmv], dtype=np.float64)
    tab = None
    if draws is not None:
        tab = np.zeros((steps, 12))
        for k in range(steps):
            for f, d in enumerate(draws[k]):
                if d is not None:
                    tab[k, 6 * f:6 * f + 5] = d
                    tab[k, 6 * f + 5] = 1.0
    rc = lib().r2d3c_cpu_run(n, _ptr(prm), _ptr(ps), _ptr(vz), int(steps), float(dt),
                             None if tab is None else _ptr(tab))
    if rc != 0:
        raise ValueError("r2d3c_cpu_run failed (%d)" % rc)
