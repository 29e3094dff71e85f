"""The fit-null workloads at reduced size against the oracle's results (committed fixtures
written by tests/golden/make_config_vectors.py from the same seeded cohorts):

  families: 400 nuclear families (n = 2,000) x 4,000 markers, 5% prevalence, one trait;
  biobank:  8,000 samples (15% half-sib pairs) x 4,000 markers, 4 traits at 0.5/1/5/20%,
            fitted together with joined GRM products.

Bars: the FP64 path matches the oracle to 1e-6 in tau and VR with the same AI-REML
iterations, and to 1e-6 in step-2 -log10 p; the reduced-precision forms to 1e-4 in tau and
VR, 1e-3 in -log10 p, iterations within one."""
import os

import numpy as np
import pytest

import csaige_amd as sg
from csaige_amd import synth
from tests.helpers import neglog10, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FORMS = [sg.PREC_FP64, sg.PREC_INT8, sg.PREC_INT8X2, sg.PREC_MIXED]
IDS = ["fp64", "int8", "int8x2", "bf16"]


def _gold(name):
    import tests.golden.make_config_vectors as mk
    g = np.load(os.path.join(GOLD, name))
    return mk, {k: g[k] for k in g.files}


def _pvalues(fit, y, c, q):
    e = sg.ScoreTestEngine(fit["mu"], y - fit["mu"], fit["tau"], c.x, fit["xvx_inv_xv"],
                           fit["vr"], y=y)
    r = e.test_markers_u8(q)
    e.close()
    return r


def _check(prec, got, want_tau, want_it, want_vr, want_p, pv, label):
    dtau = rel(got["tau"], want_tau).max()
    dvr = abs(got["vr"] - want_vr) / want_vr
    m = want_p >= 1e-10
    dp = np.abs(neglog10(pv["pvalue"][m]) - neglog10(want_p[m])).max()
    print(f"{label} prec {prec}: tau {got['tau'][1]:.8g} (oracle {want_tau[1]:.8g}, rel "
          f"{dtau:.2e}) it {got['iterations']}/{int(want_it)} vr rel {dvr:.2e} "
          f"max |d -log10 p| {dp:.2e}", flush=True)
    assert got["converged"]
    if prec == sg.PREC_FP64:
        assert dtau < 1e-6 and dvr < 1e-6 and dp < 1e-6
        assert got["iterations"] == int(want_it)
    else:
        assert dtau < 1e-4 and dvr < 1e-4 and dp < 1e-3
        assert abs(got["iterations"] - int(want_it)) <= 1


@pytest.mark.parametrize("prec", FORMS, ids=IDS)
def test_families_config_vs_oracle(prec):
    mk, g = _gold("config_families.npz")
    c = mk.families()
    assert mk.cohort_hash(c) == str(g["cohort_sha256"])
    idx = np.arange(c.n, dtype=np.uint32)
    got = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                      precision=prec)
    np.testing.assert_array_equal(got["marker_class"], g["marker_class"])
    pv = _pvalues(got, c.y, c, mk.test_dosages(c.n))
    _check(prec, got, g["tau"], g["iterations"], float(g["vr"]), g["pvalue"], pv, "families")


@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_INT8, sg.PREC_INT8X2],
                         ids=["fp64", "int8", "int8x2"])
def test_biobank_four_traits_joined_vs_oracle(prec):
    mk, g = _gold("config_biobank.npz")
    c = mk.biobank()
    assert mk.cohort_hash(c) == str(g["cohort_sha256"])
    assert c.y[0].sum() >= 40  # 40 cases at 0.5% prevalence
    idx = np.arange(c.n, dtype=np.uint32)
    fits = sg.fit_null_multi(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                             precision=prec)
    q = mk.test_dosages(c.n)
    for t, got in enumerate(fits):
        np.testing.assert_array_equal(got["marker_class"], g["marker_class"][t])
        pv = _pvalues(got, c.y[t], c, q)
        _check(prec, got, g["tau"][t], g["iterations"][t], float(g["vr"][t]), g["pvalue"][t],
               pv, f"biobank trait {t}")


def test_headline_cohort_joined_int8x2_vs_fp64():
    """The headline configuration at its own size: 400,000 samples (15% in half-sib pairs) x
    150,000 markers, 4 binary traits at 0.5/1/5/20% fitted together in the int8x2 form, the
    cohort and seeds of the benchmark.  Trait 0 (0.5%, the fewest cases) and trait 3 (20%, the
    most PCG iterations) are also fitted alone in the plain FP64 form, which matches the oracle
    at the reduced sizes above: tau and VR within 1e-4, the same AI-REML iterations, and
    step-2 -log10 p within 1e-3 under both null models."""
    from csaige_amd import hip
    c = synth.biobank(n=400000, m=150000, frac_related=0.15, prevalences=(0.005, 0.01, 0.05, 0.20),
                      seed=5)
    n, m = c.n, c.bed.shape[0]
    idx = np.arange(n, dtype=np.uint32)
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)
    joined = sg.fit_null_multi(None, c.bpm, n, m, c.y, c.x, sample_idx=idx,
                               precision=sg.PREC_INT8X2, d_bed=d_bed.value)
    q, _ = synth.dosage_u8(n, 0, 256, seed=8, maf=(1e-3, 0.5))
    for t in (0, 3):
        plain = sg.fit_null(None, c.bpm, n, m, c.y[t], c.x, sample_idx=idx,
                            precision=sg.PREC_FP64, d_bed=d_bed.value)
        got = joined[t]
        pv = _pvalues(got, c.y[t], c, q)
        pw = _pvalues(plain, c.y[t], c, q)
        mk = pw["pvalue"] >= 1e-10
        dp = np.abs(neglog10(pv["pvalue"][mk]) - neglog10(pw["pvalue"][mk])).max()
        dtau = rel(got["tau"], plain["tau"]).max()
        dvr = abs(got["vr"] - plain["vr"]) / plain["vr"]
        print(f"trait {t}: tau {got['tau'][1]:.8g} vs fp64 {plain['tau'][1]:.8g} (rel {dtau:.2e}), "
              f"it {got['iterations']}/{plain['iterations']}, vr rel {dvr:.2e}, "
              f"max |d -log10 p| {dp:.2e}", flush=True)
        assert got["converged"] and plain["converged"]
        np.testing.assert_array_equal(got["marker_class"], plain["marker_class"])
        assert dtau < 1e-4 and dvr < 1e-4 and dp < 1e-3
        assert got["iterations"] == plain["iterations"]
