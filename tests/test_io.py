"""CPU tests of the file formats (rot2d3c_amd.io): Fortran E/F edit
descriptors, parameter.inp / status.inp, per-rank unformatted slab dumps."""
import os
import struct

import numpy as np
import pytest

from rot2d3c_amd import io as rio

PARAM_TEXT = """  0.5d0               % 1  dt
  12                  % 2  step
  1                   % 3  cstep, balance files
  4 %10000            % 4  sstep, spectra files
  6 %200000           % 5  tstep, flow field files
  1.00d0              % 6  fp0
  0.50d0              % 7  fz0
  0.010d0             % 8  u0
  0.020d0             % 9  v0
  5.0d0               % 10 kdn
  5.0d0               % 11 kup
  0.0010d0            % 12 nu
  0.5d-1              % 13 hnu
  1                   % 14 nn
  2                   % 15 mm
  0.0020d0            % 16 nuv
  1.00000d0           % 17 hnuv
  1                   % 18 nnv
  2                   % 19 mmv
  1.50d0              % 20 omega
  7                   % 21 seed
  3                   % 22 iflow
  1.000d0             % 23 prm1
  1.000d0             % 24 prm2
  0.0000d0            % 25 dt_corr
../ins
../outs
"""


def test_parameter_inp_read_and_round_trip(tmp_path):
    p = tmp_path / "parameter.inp"
    p.write_text(PARAM_TEXT)
    q = rio.read_parameter_inp(str(p))
    assert q["CFL"] == 0.5 and q["step"] == 12 and q["sstep"] == 4 and q["tstep"] == 6
    assert q["hnu"] == 0.05 and q["nn"] == 1 and q["mm"] == 2 and q["omega"] == 1.5
    assert q["seed"] == 7 and q["iflow"] == 3 and q["idir"] == "../ins" and q["odir"] == "../outs"
    p2 = tmp_path / "p2.inp"
    rio.write_parameter_inp(str(p2), q)
    assert rio.read_parameter_inp(str(p2)) == q
    # the analysis scripts parse column 0 after translating 'd' -> '0'
    # (python_scripts/energy.py:18,51): every written value must survive that
    for line in p2.read_text().splitlines()[:25]:
        tok = line.split()[0]
        float(tok.replace("d", "0")) if "d" in tok else int(tok)


def test_status_inp(tmp_path):
    p = tmp_path / "status.inp"
    p.write_text("0           % last output file\n0.0         % time of last output\n")
    assert rio.read_status_inp(str(p)) == (0, 0.0)
    rio.write_status_inp(str(p), 3, 1.25)
    assert rio.read_status_inp(str(p)) == (3, 1.25)


@pytest.mark.parametrize("x,w,d,e,expect", [
    (1.0, 23, 14, 3, "  0.10000000000000E+001"),
    (-1.5, 23, 14, 3, " -0.15000000000000E+001"),
    (0.0, 23, 14, 3, "  0.00000000000000E+000"),
    (1.234567890123456e-5, 23, 14, 3, "  0.12345678901235E-004"),
    (9.999999999999999e2, 24, 15, 3, "  0.100000000000000E+004"),
    (999.999999999999, 24, 15, 3, "  0.999999999999999E+003"),
    (0.99999999999999999, 23, 14, 3, "  0.10000000000000E+001"),
    (123.456, 24, 15, 3, "  0.123456000000000E+003"),
])
def test_fortran_e_edit_descriptor(x, w, d, e, expect):
    s = rio.fortran_e(x, w, d, e)
    assert s == expect and len(s) == w
    assert float(s) == pytest.approx(x, rel=10.0 ** (1 - d), abs=0)


def test_fortran_f_and_time_lines():
    assert rio.fortran_f(0.0, 12, 6) == "    0.000000"
    assert rio.fortran_f(12.3456789, 12, 6) == "   12.345679"


@pytest.mark.parametrize("nprocs", [1, 3, 4])
def test_field_slabs_round_trip_and_script_reader(tmp_path, nprocs):
    n = 12
    r = np.random.default_rng(0).standard_normal((n, n))
    rio.write_field(str(tmp_path), "ps", 7, r, nprocs)
    assert np.array_equal(rio.read_field(str(tmp_path), "ps", 7, n, nprocs), r)
    # the slab split of `range` (src/fftp.f90:40-44)
    sizes = [rio.block_range(1, n, nprocs, k) for k in range(nprocs)]
    assert sizes[0][0] == 1 and sizes[-1][1] == n
    assert all(sizes[k][1] + 1 == sizes[k + 1][0] for k in range(nprocs - 1))
    # what the analysis scripts do (python_scripts/field_calc.py:13-18): skip
    # the 4-byte marker, read the slab, stack the slabs
    parts = []
    for k in range(nprocs):
        fn = os.path.join(str(tmp_path), "ps.%03d.007.out" % k)
        raw = open(fn, "rb").read()
        assert struct.unpack("<i", raw[:4])[0] == len(raw) - 8
        parts.append(np.frombuffer(raw[4:-4], dtype="d"))
    assert np.array_equal(np.concatenate(parts).reshape(n, n), r)


def test_balance_rows_load_with_numpy(tmp_path):
    p = tmp_path / "energy_bal.txt"
    rows = [[0.0, 1.5, -2e-7, 3.25e10, 0, 1, 2, 3, 4, 5, 6, 7], [0.1] * 12]
    rio.append_rows(str(p), rows)
    back = np.loadtxt(str(p))
    assert back.shape == (2, 12)
    assert np.allclose(back, rows, rtol=1e-13)
    assert all(len(l) == 12 * 23 for l in p.read_text().splitlines())
