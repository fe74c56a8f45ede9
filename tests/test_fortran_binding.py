"""The ISO_C_BINDING module a Fortran maintainer would add
(rot2d3c_amd/fortran/rot2d3c_c.f90) compiles with the image's Fortran
compiler and links against librot2d3c.so; its host-only entry points give the
oracle's values (no device work)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from oracle import spectral as O

FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROGRAM = """program t
  use rot2d3c_c
  implicit none
  integer(c_int) :: s
  real :: r
  real(c_double) :: kx, ky, ph
  type(r2d3c_forcing) :: f
  type(r2d3c_params) :: p
  type(r2d3c_force_draw) :: d
  s = 7
  r = r2d3c_randu(s)
  write(*,'(ES16.8,I12)') r, s
  s = 11
  call r2d3c_rand_force_draw(10.d0, s, kx, ky, ph)
  write(*,'(3ES25.16,I12)') kx, ky, ph, s
  write(*,'(3I8)') c_sizeof(p), c_sizeof(d), c_sizeof(f)
end program
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_binding_links_and_matches_oracle(tmp_path):
    libdir = os.path.join(ROOT, "rot2d3c_amd")
    if not os.path.exists(os.path.join(libdir, "librot2d3c.so")):
        pytest.skip("librot2d3c.so not built")
    mod = os.path.join(ROOT, "rot2d3c_amd", "fortran", "rot2d3c_c.f90")
    (tmp_path / "t.f90").write_text(PROGRAM)
    subprocess.run([FC, "-c", mod, "-o", "m.o"], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run([FC, "t.f90", "m.o", "-L" + libdir, "-lrot2d3c", "-Wl,-rpath," + libdir, "-o", "t"],
                   cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    r, s = O.randu(7)
    assert float(out[0]) == pytest.approx(float(r), rel=1e-7) and int(out[1]) == s
    kx, ky, ph, s2 = O.rand_force_draw(10.0, 11)
    assert (float(out[2]), float(out[3]), int(out[5])) == (kx, ky, s2)
    assert float(out[4]) == pytest.approx(ph, rel=1e-15)
    # the derived types have the C structs' sizes (a short type would make
    # r2d3c_step read past the caller's struct)
    import ctypes
    from rot2d3c_amd import _lib
    sizes = [ctypes.sizeof(t) for t in (_lib.Params, _lib.ForceDraw, _lib.Forcing)]
    assert [int(x) for x in out[6:9]] == sizes
