"""The HBM traffic tool (rot2d3c_amd.traffic) on synthetic counter files:
calibration by the copy probe, kernel classes (the persistent and the
non-persistent row kernels together, fused-column dispatches split by stage,
the column radix form the run used most) and the per-launch figures against
the algorithmic bytes."""
import csv

import pytest

from rot2d3c_amd import traffic as T

HEAD = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, rows, counter):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for i, (name, v) in enumerate(rows):
            w.writerow([i + 1, i + 1, name, counter, v])


def test_classes_scaling_and_ratios(tmp_path):
    per, _, _ = T.algorithmic_bytes(2048, 10)
    kib = lambda b: b / 1024.0
    # FETCH_SIZE reads half of what is fetched, WRITE_SIZE all of it
    fetch = [("k_copy_probe(...)", kib(T.COPY_BYTES) / 2), ("k_copy_probe(...)", kib(T.COPY_BYTES) / 2),
             ("void at::native::vectorized_elementwise_kernel<fill>", kib(4 * T.COPY_BYTES)),
             ("void k_step_colinv<2048>(...)", kib(100e6) / 2)]
    write = [("k_copy_probe(...)", kib(T.COPY_BYTES)), ("k_copy_probe(...)", kib(T.COPY_BYTES)),
             ("void at::native::vectorized_elementwise_kernel<fill>", kib(4 * T.COPY_BYTES)),
             ("void k_step_colinv<2048>(...)", kib(50e6))]
    for step in range(3):
        for stage in range(4):
            name = "void k_step_row<2048>(...)" if step else "void k_step_row2<2048, 2>(...)"
            fetch.append((name, kib(130e6) / 2))
            write.append((name, kib(45e6)))
            fused = 70e6 if stage < 3 else 90e6
            fetch.append(("void k_step_colfused<2048>(...)", kib(fused) / 2))
            write.append(("void k_step_colfused<2048>(...)", kib(140e6)))
    fetch.append(("void k_ens_colfused<256>(...)", 1e9))  # not a 2048 step kernel
    write.append(("void k_ens_colfused<256>(...)", 1e9))
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fp, fetch, "FETCH_SIZE")
    _write(wp, write, "WRITE_SIZE")
    res = T.build(str(fp), str(wp), 10)
    assert res["calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert res["calibration"]["write_factor"] == pytest.approx(1.0)
    k = res["kernels"]
    assert k["k_step_row"]["hbm_bytes_per_launch"] == pytest.approx(175e6, rel=1e-9)
    assert k["k_step_colfused"]["fetch_bytes_per_launch"] == pytest.approx(70e6, rel=1e-9)
    assert k["k_step_colfused_last"]["fetch_bytes_per_launch"] == pytest.approx(90e6, rel=1e-9)
    assert k["k_step_colinv"]["hbm_bytes_per_launch"] == pytest.approx(150e6, rel=1e-9)
    for name, v in k.items():
        assert v["algorithmic_bytes_per_launch"] == round(per[name])
        assert v["ratio_to_algorithmic"] == pytest.approx(v["hbm_bytes_per_launch"] / per[name],
                                                          abs=1e-4)


def test_radix_form_template_names(tmp_path):
    """Column kernels named with their radix form: the form with the most
    dispatches (here `<2048, false>`) is counted, the other form's tuning
    runs are not, and they do not shift the stage count."""
    kib = lambda b: b / 1024.0
    fetch = [("k_copy_probe(...)", kib(T.COPY_BYTES) / 2)]
    write = [("k_copy_probe(...)", kib(T.COPY_BYTES))]
    for form in ("true", "false", "false"):
        for stage in range(4):
            fused = 70e6 if stage < 3 else 90e6
            if form == "true":
                fused *= 10
            name = "void k_step_colfused<2048, %s>(...)" % form
            fetch.append((name, kib(fused) / 2))
            write.append((name, kib(140e6)))
    fetch.append(("void k_step_colfused<2048, true>(...)", 1e9))  # a lone other-form dispatch
    write.append(("void k_step_colfused<2048, true>(...)", 1e9))
    fetch.append(("void k_step_colinv<2048, false>(...)", kib(100e6) / 2))
    write.append(("void k_step_colinv<2048, false>(...)", kib(50e6)))
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fp, fetch, "FETCH_SIZE")
    _write(wp, write, "WRITE_SIZE")
    res = T.build(str(fp), str(wp), 10)
    assert res["column_form"] == "four-step"
    k = res["kernels"]
    assert k["k_step_colfused"]["fetch_bytes_per_launch"] == pytest.approx(70e6, rel=1e-9)
    assert k["k_step_colfused_last"]["fetch_bytes_per_launch"] == pytest.approx(90e6, rel=1e-9)
    assert k["k_step_colinv"]["hbm_bytes_per_launch"] == pytest.approx(150e6, rel=1e-9)


def test_missing_copy_probe_is_an_error(tmp_path):
    fp = tmp_path / "f.csv"
    _write(fp, [("void k_step_row<2048>(...)", 1.0),
                ("void at::native::vectorized_elementwise_kernel<copy>", 1e6)], "FETCH_SIZE")
    with pytest.raises(ValueError):
        T.build(str(fp), str(fp), 10)


def test_algorithmic_bytes_of_the_row_kernel():
    """8 mixed fields of n (K + 1) complex doubles, K = floor(n / 3)."""
    per, step, geo = T.algorithmic_bytes(2048, 10)
    assert geo["K"] == 682 and per["k_step_row"] == 8 * 2048 * 683 * 16
    assert geo["B76S"] == 76 * 16 * 2048 * 1025


def test_stockham_form_is_reported_when_it_ran_most(tmp_path):
    kib = lambda b: b / 1024.0
    fetch = [("k_copy_probe(...)", kib(T.COPY_BYTES) / 2)]
    write = [("k_copy_probe(...)", kib(T.COPY_BYTES))]
    for form, reps in (("false", 1), ("true", 2)):
        for _ in range(reps):
            for stage in range(4):
                fused = (70e6 if stage < 3 else 90e6) * (1 if form == "true" else 3)
                name = "void k_step_colfused<2048, %s>(...)" % form
                fetch.append((name, kib(fused) / 2))
                write.append((name, kib(140e6)))
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fp, fetch, "FETCH_SIZE")
    _write(wp, write, "WRITE_SIZE")
    res = T.build(str(fp), str(wp), 10)
    assert res["column_form"] == "radix-8 Stockham"
    assert res["kernels"]["k_step_colfused"]["fetch_bytes_per_launch"] == pytest.approx(70e6, rel=1e-9)
    assert res["kernels"]["k_step_colfused_last"]["fetch_bytes_per_launch"] == pytest.approx(90e6, rel=1e-9)
