"""HBM traffic per launch of the 2048^2 step kernels from two rocprofv3 counter
runs, and the algorithmic (compulsory) bytes they are judged against.

Usage (FETCH_SIZE and WRITE_SIZE each from its own run of
`python bench.py --steps 10 --warmup 2 --cpu-steps 0 --ensemble-steps 0
--no-closure --large-steps 0` under `rocprofv3 --pmc <counter> --output-format
csv`):

    python -m rot2d3c_amd.traffic FETCH.csv WRITE.csv --steps 10 --out traffic.json

Each counter is scaled by what it reads for the bench's streaming copy probe
(`k_copy_probe`, 1 GiB each way per call): FETCH_SIZE reads about half the
bytes fetched on this GPU, WRITE_SIZE all the bytes written.  The column
kernels come in two radix forms; the form with the most dispatches (the one
the bench's schedule used) is reported.
"""
import argparse
import csv
import json
import sys

import numpy as np

COPY_BYTES = 1 << 30  # bytes each way per copy-probe call


def algorithmic_bytes(n, k):
    """Compulsory HBM bytes per launch of each step kernel class for a run of k
    chained steps (one r2d3c_run call), dealiased schedule.

    M = bytes of one mixed field restricted to the kx <= K columns every
    dealiased field lives in, D' = number of (kx, ky) half-plane modes inside
    the update mask tiny < k^2 <= kmax (D = D' + 1 with the k = 0 mode),
    S = bytes of one full spectral field.
      column inverse (once per run): read psi, vz on the disc, write 6 M
      row kernel: read 6 M, write 2 M
      fused column, stages 1-3: read 2 M + psi0, vz0 on the disc, write 6 M
      fused column, last stage: read 2 M + psi0, vz0, write psi, vz on the disc
        and, except after the run's last step, the next step's 6 M."""
    kmax = (n / 3.0) ** 2
    K = 0
    while (K + 1) ** 2 <= kmax:
        K += 1
    M = n * (K + 1) * 16
    ky = np.fft.fftfreq(n, 1.0 / n)
    kx = np.arange(K + 1, dtype=np.float64)
    D = int(np.count_nonzero(kx[:, None] ** 2 + ky[None, :] ** 2 <= kmax))
    S = 16 * n * (n // 2 + 1)
    disc = 2 * 16 * (D - 1)
    last_chained = 2 * M + disc + disc + 6 * M
    last_final = 2 * M + disc + disc
    per = {
        "k_step_colinv": 2 * 16 * D + 6 * M,
        "k_step_row": 8 * M,
        "k_step_colfused": 2 * M + disc + 6 * M,
        "k_step_colfused_last": ((k - 1) * last_chained + last_final) / k,
    }
    step = per["k_step_colinv"] / k + 4 * per["k_step_row"] + 3 * per["k_step_colfused"] \
        + per["k_step_colfused_last"]
    return per, step, {"K": K, "M": M, "D": D, "S": S, "B76S": 76 * S}



def _dispatches(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in rows]


def _copy_reading(disp):
    """The counter's median reading (KiB) over the copy-probe dispatches."""
    vals = [v for name, v in disp if "k_copy_probe" in name]
    if not vals:
        raise ValueError("no k_copy_probe dispatch in the counter run (bench.py runs one)")
    return float(np.median(vals))


def _column(name, kern, form):
    """A 2048 column kernel of the given radix form ("false": four-step,
    "true": radix-8 Stockham; a name without the form is the four-step)."""
    if form == "false" and (kern + "<2048>") in name:
        return True
    return (kern + "<2048, %s>" % form) in name


def _form(disp):
    """The column form with the most fused-column dispatches."""
    n = {f: sum(1 for name, _ in disp if _column(name, "k_step_colfused", f))
         for f in ("false", "true")}
    return max(n, key=lambda f: n[f])


def _classes(disp, form):
    """Readings (KiB) per step-kernel class; fused-column dispatches are
    stages 4, 3, 2, 1 in turn, the fourth of each step being the last stage."""
    out = {"k_step_colinv": [], "k_step_row": [], "k_step_colfused": [], "k_step_colfused_last": []}
    fused = 0
    for name, v in disp:
        if _column(name, "k_step_colinv", form):
            out["k_step_colinv"].append(v)
        elif "k_step_row<2048" in name or "k_step_row2<2048" in name:
            out["k_step_row"].append(v)
        elif _column(name, "k_step_colfused", form):
            out["k_step_colfused_last" if fused % 4 == 3 else "k_step_colfused"].append(v)
            fused += 1
    return out


def build(fetch_csv, write_csv, steps, n=2048):
    fd, wd = _dispatches(fetch_csv), _dispatches(write_csv)
    ff = COPY_BYTES / (_copy_reading(fd) * 1024.0)
    wf = COPY_BYTES / (_copy_reading(wd) * 1024.0)
    form = _form(fd)
    fc, wc = _classes(fd, form), _classes(wd, form)
    per, _, _ = algorithmic_bytes(n, steps)
    kern = {}
    for k in fc:
        if not fc[k] or not wc[k]:
            continue
        fb = float(np.median(fc[k])) * 1024.0 * ff
        wb = float(np.median(wc[k])) * 1024.0 * wf
        kern[k] = {"fetch_bytes_per_launch": int(round(fb)), "write_bytes_per_launch": int(round(wb)),
                   "hbm_bytes_per_launch": int(round(fb + wb)),
                   "algorithmic_bytes_per_launch": int(round(per[k])),
                   "ratio_to_algorithmic": round((fb + wb) / per[k], 4)}
    return {"workload": "%d^2 FP64 RK4 step (bench.py --steps %d --warmup 2 --cpu-steps 0 "
                        "--ensemble-steps 0 --no-closure --large-steps 0)" % (n, steps),
            "column_form": "radix-8 Stockham" if form == "true" else "four-step",
            "calibration": {"copy_bytes": COPY_BYTES, "FETCH_SIZE_KiB": _copy_reading(fd),
                            "WRITE_SIZE_KiB": _copy_reading(wd), "fetch_factor": ff,
                            "write_factor": wf},
            "kernels": kern}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    res = build(a.fetch_csv, a.write_csv, a.steps)
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    for k, v in res["kernels"].items():
        print(k, v["hbm_bytes_per_launch"], v["ratio_to_algorithmic"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
