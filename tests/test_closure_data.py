"""The committed trained closure network and its evaluation record are well
formed: parameter count and layout of the library, padding exactly zero,
finite values, and a record whose numbers are consistent."""
import json
import os

import numpy as np

from conftest import ROOT
from oracle import cnn as C

DATA = os.path.join(ROOT, "rot2d3c_amd", "data")


def test_trained_parameters_have_the_library_layout():
    z = np.load(os.path.join(DATA, "closure_256.npz"))
    p = z["params"]
    offs, total = C.offsets()
    assert p.dtype == np.float32 and p.shape == (total,)
    assert np.all(np.isfinite(p))
    W0 = p[offs[0][0]:offs[0][1]].reshape(9, 32, 16)
    W5 = p[offs[5][0]:offs[5][1]].reshape(9, 16, 32)
    assert np.all(W0[:, :, 3:] == 0) and np.all(W5[:, 2:, :] == 0)
    assert np.all(p[offs[5][1] + 2:offs[5][1] + 16] == 0)
    assert z["scale_in"].shape == (3,) and z["scale_out"].shape == (2,)
    assert np.all(z["scale_in"] > 0) and np.all(z["scale_out"] > 0)
    assert int(z["nf"]) == 2048 and int(z["nc"]) == 256
    # inference settings chosen on the tuning window
    assert float(z["out_clamp"]) >= 0.0 and 0.0 <= float(z["sign_clip"]) <= 1.0
    on, off = (float(v) for v in z["safe_cfl"])
    assert 0.0 < off <= on


def test_evaluation_record_is_consistent():
    ev = json.load(open(os.path.join(DATA, "closure_256_eval.json")))
    rc = np.array(ev["ratio_closure_to_dns"])
    assert len(rc) == ev["shells"] == 85
    assert np.isclose(np.abs(rc - 1).max(), ev["closure"]["spectrum_rel_err_max"], atol=1e-3)
    assert np.isclose(np.abs(rc - 1).mean(), ev["closure"]["spectrum_rel_err_mean"], atol=1e-3)
    assert ev["shells_within_5pct"] == int(np.sum(np.abs(rc - 1) <= 0.05))
    # the closure run is better than the plain coarse run everywhere it matters
    assert ev["closure"]["spectrum_rel_err_mean"] < ev["plain"]["spectrum_rel_err_mean"]
    assert ev["closure"]["flux_err"] < ev["plain"]["flux_err"]
    # the recorded evaluation: stable (no fallback to the plain solver) and
    # within 5% on average over the shells k <= 85
    assert ev["closure"]["fallbacks"] == 0
    assert 0.0 <= ev["closure"]["safe_fraction"] < 0.5
    assert ev["closure"]["spectrum_rel_err_mean"] <= 0.05


def test_evaluation_judge_is_consistent():
    """The judge against the mean of DNS realisations forced by the coarse
    runs' draws: a shell agrees when its relative error is within 5% or
    within the noise floor (2 sigma of the difference of the two means)."""
    ev = json.load(open(os.path.join(DATA, "closure_256_eval.json")))
    for name in ("closure", "plain"):
        j = ev[name]["judge"]
        r, nz = np.array(j["ratio"]), np.array(j["noise_floor_rel"])
        assert len(r) == len(nz) == j["shells"] == 85
        ok = np.abs(r - 1) <= np.maximum(0.05, nz)
        assert j["shells_within_5pct_or_noise"] == int(ok.sum())
        assert j["shells_within_5pct"] == int(np.sum(np.abs(r - 1) <= 0.05))
        assert j["shells_outside"] == [int(k) + 1 for k in np.nonzero(~ok)[0]]
        assert j["dns_realisations"] >= 2
    assert (ev["closure"]["judge"]["shells_within_5pct_or_noise"] >
            ev["plain"]["judge"]["shells_within_5pct_or_noise"])
