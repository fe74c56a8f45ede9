"""Input and output files of a Rot2D3C run, in the reference's formats.

  parameter.inp  25 list-directed values + input and output directories
                 (read order src/Rot2D3C.para.f90:136-162)
  status.inp     last output index and its time (src/Rot2D3C.para.f90:125-129)
  <name>.<rank3>.<out3>.out
                 real-space field slabs, one Fortran sequential unformatted
                 record per rank: 4-byte length, n*(jend-jsta+1) float64 with x
                 fastest, 4-byte length (src/Rot2D3C.para.f90:483-518); the
                 slab ranges follow `range` (src/fftp.f90:17-47)
  energy_bal.txt, enstrophy_bal.txt, helicity_bal.txt, max_min.txt
                 FORMAT E23.14E3 columns (src/pseudospec2D_Rot2D3C.f90:592-608)
  spectrum.XXXX.txt, transfer.XXXX.txt, fluxes.XXXX.txt
                 FORMAT E24.15E3 columns (src/pseudospec2D_Rot2D3C.f90:699-709,815-834)
  time_field.txt (A3,F12.6), time_spec.txt (A4,F12.6)
"""
import math
import os
import struct

import numpy as np

PARAM_NAMES = ("CFL", "step", "cstep", "sstep", "tstep", "fp0", "fz0", "u0", "v0", "kdn", "kup",
               "nu", "hnu", "nn", "mm", "nuv", "hnuv", "nnv", "mmv", "omega", "seed", "iflow",
               "prm1", "prm2", "dt_corr")
INT_PARAMS = {"step", "cstep", "sstep", "tstep", "nn", "mm", "nnv", "mmv", "seed", "iflow"}


def _first_value(line):
    """First item of a list-directed record; `d` exponents as in Fortran."""
    tok = line.replace(",", " ").split()
    if not tok:
        raise ValueError("empty record in parameter file")
    return tok[0]


def _fnum(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def read_parameter_inp(path):
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < 27:
        raise ValueError("%s: expected 27 records, found %d" % (path, len(lines)))
    p = {}
    for k, name in enumerate(PARAM_NAMES):
        tok = _first_value(lines[k])
        p[name] = int(_fnum(tok)) if name in INT_PARAMS else _fnum(tok)
    p["idir"] = lines[25][:100].strip()
    p["odir"] = lines[26][:100].strip()
    return p


def _dnum(x):
    """A real written with a `d0`-style exponent, as the python_scripts expect
    (they translate `d` to `0` before parsing: 1.0d-3 -> 1.00-3 would fail, so
    values are written as mantissa `d0` when possible)."""
    x = float(x)
    s = repr(x)
    if "e" in s or "E" in s:
        m, e = ("%.15e" % x).split("e")
        m = m.rstrip("0").rstrip(".") if "." in m else m
        return "%sd%d" % (m, int(e))
    if "." not in s:
        s += ".0"
    return s + "d0"


def write_parameter_inp(path, p):
    with open(path, "w") as f:
        for k, name in enumerate(PARAM_NAMES):
            v = p[name]
            val = str(int(v)) if name in INT_PARAMS else _dnum(v)
            f.write("%-20s%% %d %s\n" % (val, k + 1, name))
        f.write(p["idir"] + "\n")
        f.write(p["odir"] + "\n")


def read_status_inp(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return int(_fnum(_first_value(lines[0]))), _fnum(_first_value(lines[1]))


def write_status_inp(path, stat, time):
    with open(path, "w") as f:
        f.write("%d           %% last output file\n%s         %% time of last output\n"
                % (stat, repr(float(time))))


def block_range(n1, n2, nprocs, irank):
    """Even split of [n1, n2] over nprocs ranks (src/fftp.f90:17-47); 1-based."""
    iwork1 = (n2 - n1 + 1) // nprocs
    iwork2 = (n2 - n1 + 1) % nprocs
    ista = irank * iwork1 + n1 + min(irank, iwork2)
    iend = ista + iwork1 - 1
    if iwork2 > irank:
        iend += 1
    return ista, iend


def _ext3(k):
    return "%03d" % k


def field_path(d, name, rank, out):
    return os.path.join(d, "%s.%s.%s.out" % (name, _ext3(rank), _ext3(out)))


def write_field(d, name, out, real, nprocs=1):
    """One slab file per rank of the real field real[y][x] (float64, n x n)."""
    n = real.shape[0]
    for r in range(nprocs):
        jsta, jend = block_range(1, n, nprocs, r)
        data = np.ascontiguousarray(real[jsta - 1:jend], dtype="<f8").tobytes()
        if len(data) > 2 ** 31 - 1:
            raise ValueError("record of %d bytes exceeds a 4-byte record marker: use more slabs" % len(data))
        with open(field_path(d, name, r, out), "wb") as f:
            mark = struct.pack("<i", len(data))
            f.write(mark)
            f.write(data)
            f.write(mark)


def read_field(d, name, out, n, nprocs=1):
    real = np.empty((n, n))
    for r in range(nprocs):
        jsta, jend = block_range(1, n, nprocs, r)
        with open(field_path(d, name, r, out), "rb") as f:
            (ln,) = struct.unpack("<i", f.read(4))
            if ln != 8 * n * (jend - jsta + 1):
                raise ValueError("%s: record of %d bytes, expected %d"
                                 % (field_path(d, name, r, out), ln, 8 * n * (jend - jsta + 1)))
            real[jsta - 1:jend] = np.frombuffer(f.read(ln), dtype="<f8").reshape(jend - jsta + 1, n)
            (ln2,) = struct.unpack("<i", f.read(4))
            if ln2 != ln:
                raise ValueError("record markers differ")
    return real


def fortran_e(x, w, d, e):
    """Fortran Ew.dEe edit descriptor (0.ddd..E+eee form, right-justified)."""
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w)
    if x == 0.0:
        mant, ex = "0" * d, 0
    else:
        s = "%.*e" % (d - 1, abs(x))  # m.mmm e+XX with d significant digits
        m, ee = s.split("e")
        mant, ex = m.replace(".", ""), int(ee) + 1
    sign = "-" if x < 0 or (x == 0.0 and math.copysign(1.0, x) < 0) else ""
    out = "%s0.%sE%s%0*d" % (sign, mant, "-" if ex < 0 else "+", e, abs(ex))
    return out.rjust(w)


def e23(row):
    return "".join(fortran_e(v, 23, 14, 3) for v in row)


def e24(row):
    return "".join(fortran_e(v, 24, 15, 3) for v in row)


def fortran_f(x, w, d):
    return ("%.*f" % (d, float(x))).rjust(w)


def append_rows(path, rows, fmt=e23):
    with open(path, "a") as f:
        for r in rows:
            f.write(fmt(r) + "\n")


def write_columns(path, cols, fmt=e24):
    with open(path, "w") as f:
        for r in zip(*cols):
            f.write(fmt(r) + "\n")
