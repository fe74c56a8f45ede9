"""CPU checks of the C ABI: the library loads, exports every symbol declared in
include/rot2d3c.h, and its host-side code (random numbers, argument checks)
agrees with the oracle.  No device work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import spectral as O


def _declared():
    txt = open(os.path.join(ROOT, "include", "rot2d3c.h")).read()
    return sorted(set(re.findall(r"\b(r2d3c_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from rot2d3c_amd import _lib
    assert _declared() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from rot2d3c_amd import _lib
    l = _lib.lib()
    for name in _declared():
        assert hasattr(l, name), name


def test_randu_matches_oracle():
    from rot2d3c_amd import spectral as S
    for seed in (0, 1, 2, 12345, 2**30):
        a, b = seed, seed
        for _ in range(2000):
            va, a = S.randu(a)
            vb, b = O.randu(b)
            assert a == b
            assert va == float(vb)


def test_rand_force_draw_matches_oracle():
    from rot2d3c_amd import spectral as S
    for kup in (4.0, 10.0, 40.0, 12.0):
        s1 = s2 = 5
        for _ in range(500):
            kx, ky, ph, s1 = S.rand_force_draw(kup, s1)
            okx, oky, oph, s2 = O.rand_force_draw(kup, s2)
            assert (kx, ky, ph, s1) == (okx, oky, oph, s2)


def test_argument_errors_are_reported():
    from rot2d3c_amd import _lib
    l = _lib.lib()
    h = ctypes.c_void_p()
    assert l.r2d3c_plan_create(ctypes.byref(h), 100) == -2   # not a power of two
    assert b"power of two" in l.r2d3c_last_error()
    assert l.r2d3c_plan_create(None, 64) == -1
    assert l.r2d3c_c2r(None, None, None, None) == -1
    assert l.r2d3c_plan_destroy(None) == 0


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes' Params, ForceDraw and Forcing have the sizes and field offsets
    that a C compiler gives the header's structs."""
    import ctypes
    import shutil
    import subprocess
    from rot2d3c_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"r2d3c_params": _lib.Params, "r2d3c_force_draw": _lib.ForceDraw,
               "r2d3c_forcing": _lib.Forcing, "r2d3c_member": _lib.Member}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rot2d3c.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        for fname, _ in ct._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, fname))
    lines.append("return 0; }")
    (tmp_path / "l.c").write_text("\n".join(lines))
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "l.c", "-o", "l"], cwd=tmp_path,
                   check=True, capture_output=True)
    got = [int(x) for x in subprocess.run(["./l"], cwd=tmp_path, check=True, capture_output=True,
                                          text=True).stdout.split()]
    expect = []
    for ct in structs.values():
        expect.append(ctypes.sizeof(ct))
        expect += [getattr(ct, f).offset for f, _ in ct._fields_]
    assert got == expect

