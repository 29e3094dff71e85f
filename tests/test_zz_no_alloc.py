"""saige_test_markers_u8_device_noalloc against saige_test_markers_u8_device through the C ABI:
the same results on two calls in a row, the second with a smaller batch, for the
intercept-only engine (byte-table SPA passes) and a two-covariate one, at 70,000 samples
(above the size at which the allocating entry point tunes its kernel shape)."""
import ctypes as C

import numpy as np
import pytest

import csaige_amd as sg
from csaige_amd import hip, synth

pytestmark = pytest.mark.gpu

INT_FIELDS = ("n", "is_spa", "spa_converged", "n_cases", "n_controls")
EXACT_FIELDS = ("ac", "af", "mac", "af_cases", "af_controls", "variance_ratio")
FLOAT_FIELDS = ("score", "var_t", "var_t_star", "tstat", "beta", "se_beta")


def _results(d_r, nvar):
    rsize = C.sizeof(sg._lib.VariantResult)
    host = np.empty(nvar * rsize, np.uint8)
    d_r.download(host)
    return sg.results_to_dict((sg._lib.VariantResult * nvar).from_buffer(host))


def _compare(got, want):
    for f in INT_FIELDS + EXACT_FIELDS:
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    for f in FLOAT_FIELDS:
        np.testing.assert_allclose(got[f], want[f], rtol=1e-12, atol=1e-300, err_msg=f)
    lg = -np.log10(np.clip(got["pvalue"], 1e-300, None))
    lw = -np.log10(np.clip(want["pvalue"], 1e-300, None))
    assert np.abs(lg - lw).max() < 1e-9


@pytest.mark.parametrize("p", [1, 2], ids=["intercept", "two-covariates"])
def test_noalloc_matches_allocating_entry_point(p):
    n = 70000
    nf = synth.null_fit(n=n, p=p, eta_mean=-3.5, eta_sd=0.6, seed=11 + p)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    eng = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                             spa_pval_cutoff=0.0455)
    rsize = C.sizeof(sg._lib.VariantResult)
    B = 4096
    q, _ = synth.dosage_u8(n, 0, B, seed=21, maf=(5e-4, 0.5))
    d_q = hip.DeviceBuffer(q.nbytes)
    d_q.upload(q)
    d_a = hip.DeviceBuffer(B * rsize)
    d_b = hip.DeviceBuffer(B * rsize)
    for nvar, off in ((B, 0), (1000, 123)):  # the second call smaller, from another offset
        src = d_q.value + off * n
        eng.test_markers_u8_device(src, n, nvar, 1.0 / 127, d_a.value)
        eng.test_markers_u8_device_noalloc(src, n, nvar, 1.0 / 127, d_b.value)
        hip.device_synchronize()
        want, got = _results(d_a, nvar), _results(d_b, nvar)
        print(f"p={p} nvar={nvar}: {int(got['is_spa'].sum())} SPA'd", flush=True)
        assert got["is_spa"].sum() > 10
        _compare(got, want)
    with pytest.raises(sg.SaigeError):
        eng.test_markers_u8_device_noalloc(d_q.value, n, B + 1, 1.0 / 127, d_b.value)
    eng.close()
