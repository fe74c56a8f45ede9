"""GPU: the trained closure (rot2d3c_amd/data/closure_256.npz) judged on a
2048^2 DNS window generated in the test.  The DNS of the closure study (same
parameters and seeded start) is spun up 15 000 steps; two DNS realisations
then run 10 000 steps (2 time units) from that state, forced by the same
draws as the two realisations of each 256^2 coarse run (plain solver, and
the trained closure with its stored settings, run-time stabiliser and
fallback guards); the coarse-grained DNS spectra and fluxes, averaged in
time and over the realisations, are the judge (energy spectrum and 2D
energy flux for k <= 85; the judge arithmetic is the library's spectrum and
transfers, src/pseudospec2D_Rot2D3C.f90:615-837).

What the trained network gains over the method without it: the plain 256^2
run piles energy up towards the grid cut-off (hundreds of percent of
spectrum error, 51% flux error); the closure run stays within 5% on
average, its flux within 8% (5.9% measured, against a 3.3% noise floor of
this short window), and at least 77 of the 85 shells agree with the DNS
within 5% or within the noise floor (2 sigma of the difference of the means,
from the realisations' spread)."""
import os

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_trained_closure_against_live_dns_window(cuda):
    print("closure judge: live DNS window", flush=True)
    from rot2d3c_amd import closure_study as CS
    npz = os.path.join(ROOT, "rot2d3c_amd", "data", "closure_256.npz")
    out = CS.live_evaluation(npz, cuda, spinup=15000, window=10000, realisations=2, samples=100,
                             log=lambda *a: print(*a, flush=True))
    plain, clos = out["plain"], out["closure"]
    assert plain["finite"] and clos["finite"]
    assert clos["fallbacks"] == 0
    assert clos["spectrum_rel_err_mean"] <= 0.05, clos["spectrum_rel_err_mean"]
    print("closure flux err %.4f, its noise floor %.4f" % (clos["flux_err"],
                                                            clos["flux_noise_floor"]), flush=True)
    # measured 5.9% on this 2-unit window (noise floor 3.3%), 3.3% over the
    # study's 10-unit evaluation window with 4 realisations
    assert clos["flux_err"] <= max(0.08, clos["flux_noise_floor"]), clos["flux_err"]
    assert clos["spectrum_rel_err_mean"] < plain["spectrum_rel_err_mean"] / 20
    assert clos["flux_err"] < plain["flux_err"] / 4
    assert clos["shells"] == 85
    assert clos["shells_within_5pct_or_noise"] >= 77, clos["shells_outside"]
    assert clos["shells_within_5pct_or_noise"] > plain["shells_within_5pct_or_noise"] + 50
