"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded inputs.

Bars: integer work (decode, allele counts, MAC, marker split) bit-exact; fp64 GRM products
to 1e-12 relative; PCG iterate-identical (same iteration counts) with x to 1e-9; tau and
the variance ratio to 1e-6 relative; step-2 -log10 p to 1e-6 on the parity mask p >= 1e-10.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import csaige_amd as sg
import oracle
from csaige_amd import synth
from tests.helpers import decode_np, neglog10, random_bed, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


_ORACLE_FITS = {}


def _oracle_fit(key, bed, bpm, n, y, x, **kw):
    """oracle.fit_null, computed once per key in this session (the parametrized cases share
    their inputs)."""
    if key not in _ORACLE_FITS:
        idx = np.arange(n, dtype=np.uint32)
        _ORACLE_FITS[key] = oracle.fit_null(bed, bpm, n, bed.shape[0], idx, y, x, **kw)
    return _ORACLE_FITS[key]


# --------------------------------------------------------------------------- decode / AF
def test_marker_stats_golden_frq():
    f = np.load(os.path.join(GOLD, "frq_subset.npz"))
    n = int(f["n"])
    ac, nv, af, mac = sg.marker_stats(f["bed"], f["bed"].shape[1], n, len(f["bed"]))
    np.testing.assert_array_equal(ac, np.round(f["maf"] * f["nchrobs"]))
    np.testing.assert_array_equal(nv * 2, f["nchrobs"])
    np.testing.assert_array_equal(mac, np.minimum(ac, 2.0 * nv - ac))


def test_marker_stats_golden_r_saige_ac_subset():
    d = np.load(os.path.join(GOLD, "ac100.npz"))
    ac, nv, af, mac = sg.marker_stats(d["bed"], d["bed"].shape[1], int(d["n_total"]), len(d["bed"]),
                                      sample_idx=d["sample_idx"])
    np.testing.assert_array_equal(ac, d["ac"])
    assert (nv == 1000).all()


@pytest.mark.parametrize("n,subset", [(37, False), (1031, True), (4, False)])
def test_decode_and_stats_bitexact(n, subset):
    body, bpm = random_bed(n, 9, seed=n, p_missing=0.1)
    idx = np.random.default_rng(n).permutation(n)[: max(1, n - 5)].astype(np.uint32) if subset else None
    got = sg.decode(body, bpm, n, 9, sample_idx=idx)
    for j in range(9):
        want = oracle.decode_marker(body[j], n, idx)
        np.testing.assert_array_equal(got[j], want)
    ac, nv, af, mac = sg.marker_stats(body, bpm, n, 9, sample_idx=idx)
    for j in range(9):
        s, a, m_, k = oracle.compute_af(oracle.decode_marker(body[j], n, idx))
        assert (ac[j], nv[j], af[j], mac[j]) == (s, k, a, m_)


# --------------------------------------------------------------------------- GRM product
def _grm_pair(n, m, seed, p_missing=0.03, mono=False):
    body, bpm = random_bed(n, m, seed, p_missing)
    if mono:
        body[m // 2] = 0xFF  # all 0 dosage -> af 0 -> monomorphic
    d = decode_np(body, n)
    af = np.array([oracle.compute_af(r)[1] for r in d])
    return body, bpm, d, af


@pytest.mark.parametrize("n,m,k", [(600, 777, 1), (1000, 300, 5), (513, 1100, 34), (300, 257, 70)])
def test_grm_fp64_vs_oracle(n, m, k):
    _say(f"grm n={n} m={m} k={k}")
    body, bpm, d, af = _grm_pair(n, m, seed=n + m, mono=True)
    g = sg.OnTheFlyGrm(body, bpm, n, af)
    ref = oracle.Grm(d, af, oracle.GRM_F64)
    V = np.random.default_rng(k).standard_normal((k, n))
    got = g.mat_mat(V)
    for c in range(k):
        want = ref.mat_vec(V[c])
        np.testing.assert_allclose(got[c], want, rtol=0, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(g.mat_vec(V[0]), got[0], rtol=0, atol=0)  # deterministic


@pytest.mark.parametrize("n,m,k", [(2000, 4095, 3), (517, 1031, 34), (16, 40, 2), (1, 7, 1)])
def test_grm_f32_bit_identical_to_reference_arithmetic(n, m, k):
    """SAIGE_PREC_F32 reproduces saige-rs's f32 OnTheFlyGrm::mat_vec (pcg.rs:128-180, one rayon
    thread) bit for bit: f32 standardized values (missing -> 0, monomorphic -> 0), f32 dot in
    sample order, f32 accumulation in marker order, scale 1/M in f32."""
    _say(f"f32 grm n={n} m={m} k={k}")
    body, bpm, d, af = _grm_pair(n, m, seed=7 * n + m, p_missing=0.02, mono=m > 2)
    g = sg.OnTheFlyGrm(body, bpm, n, af, precision=sg.PREC_F32)
    ref = oracle.Grm(d, af, oracle.GRM_F32)
    V = np.random.default_rng(n + k).standard_normal((k, n)) * 3.0
    got = g.mat_mat(V)
    for c in range(k):
        np.testing.assert_array_equal(got[c], ref.mat_vec(V[c]))


def test_grm_subset_and_marker_selection():
    n_total, m_total = 700, 400
    body, bpm = random_bed(n_total, m_total, seed=5, p_missing=0.02)
    idx = np.sort(np.random.default_rng(1).choice(n_total, 650, replace=False)).astype(np.uint32)[::-1].copy()
    markers = np.arange(3, m_total, 3, dtype=np.uint32)
    d = np.stack([oracle.decode_marker(body[j], n_total, idx) for j in markers])
    af = np.array([oracle.compute_af(r)[1] for r in d])
    g = sg.OnTheFlyGrm(body, bpm, n_total, af, sample_idx=idx, marker_idx=markers)
    v = np.random.default_rng(2).standard_normal(len(idx))
    want = oracle.Grm(d, af, oracle.GRM_F64).mat_vec(v)
    np.testing.assert_allclose(g.mat_vec(v), want, atol=1e-12 * np.abs(want).max(), rtol=0)


@pytest.mark.parametrize("n_total,n,prec", [(270000, 900, None), (270000, 900, "int8"),
                                            (5003, 5003, "int8"), (1021, 700, "int8")],
                         ids=["long-rows-gather", "long-rows-gather-int8", "identity-int8",
                              "staged-list-int8"])
def test_grm_repack_paths(n_total, n, prec):
    """The three repack forms: identity sample order (funnel-shift copy), a sample list with
    rows staged in LDS, and a list over rows too long to stage (270,000 samples, 67.5 KB per
    marker: direct gather); unaligned row starts (odd bytes per marker), markers selected and
    reordered.  FP64 and the int8 path (blocked genotype values, bit-transposed sample-major
    copy)."""
    m_total = 37
    body, bpm = random_bed(n_total, m_total, seed=n_total % 97, p_missing=0.0)
    rng = np.random.default_rng(n)
    idx = None if n == n_total else rng.choice(n_total, n, replace=False).astype(np.uint32)
    markers = rng.permutation(m_total)[:29].astype(np.uint32)
    d = np.stack([oracle.decode_marker(body[j], n_total, idx) for j in markers])
    af = np.array([oracle.compute_af(r)[1] for r in d])
    kw = {} if prec is None else {"precision": sg.PREC_INT8}
    g = sg.OnTheFlyGrm(body, bpm, n_total, af, sample_idx=idx, marker_idx=markers, **kw)
    V = rng.standard_normal((3, n))
    got = g.mat_mat(V)
    ref = oracle.Grm(d, af, oracle.GRM_F64)
    for c in range(3):
        want = ref.mat_vec(V[c])
        tol = 1e-12 if prec is None else 1e-6
        assert np.linalg.norm(got[c] - want) <= tol * np.linalg.norm(want), (c, prec)


def test_grm_symmetric():
    n = 200
    body, bpm, d, af = _grm_pair(n, 300, seed=9)
    g = sg.OnTheFlyGrm(body, bpm, n, af)
    K = g.mat_mat(np.eye(n))
    np.testing.assert_allclose(K, K.T, atol=1e-13)


# --------------------------------------------------------------------------- PCG
def test_pcg_block_iterate_identical():
    n, m = 800, 1200
    body, bpm, d, af = _grm_pair(n, m, seed=17, p_missing=0)
    g = sg.OnTheFlyGrm(body, bpm, n, af)
    ref = oracle.Grm(d, af, oracle.GRM_F64)
    rng = np.random.default_rng(3)
    w = rng.uniform(0.02, 0.25, n)
    tau = (1.0, 0.8)
    B = rng.standard_normal((6, n))
    B[2] = 0.0  # converges at iteration 0
    X, its, res, conv = sg.solve_sigma(g, w, tau, B, 1e-5, 500)
    for c in range(6):
        x, it, r, cv = oracle.pcg_solve_sigma(ref, w, tau, B[c], 1e-5, 500)
        assert its[c] == it and conv[c] == cv, (c, its[c], it)
        np.testing.assert_allclose(X[c], x, rtol=0, atol=1e-9 * max(np.abs(x).max(), 1e-30))
    # max_iter stop: not converged, iterations = max_iter
    X, its, res, conv = sg.solve_sigma(g, w, tau, B[:1], 1e-12, 3)
    x, it, r, cv = oracle.pcg_solve_sigma(ref, w, tau, B[0], 1e-12, 3)
    assert its[0] == it == 3 and not conv[0] and not cv
    np.testing.assert_allclose(X[0], x, atol=1e-10 * np.abs(x).max(), rtol=0)


# --------------------------------------------------------------------------- AI-REML
@pytest.mark.parametrize("trait", [sg.TRAIT_BINARY, sg.TRAIT_QUANTITATIVE])
def test_fit_ai_reml_vs_oracle(trait):
    c = synth.unrelated(n=500, m_grm=1500, m_test=0, seed=7, h2=0.5, prevalence=0.2)
    d = decode_np(c.bed, c.n)
    af = np.array([oracle.compute_af(r)[1] for r in d])
    y = c.y if trait == sg.TRAIT_BINARY else c.liability
    g = sg.OnTheFlyGrm(c.bed, c.bpm, c.n, af)
    got = sg.fit_ai_reml(g, y, c.x, trait=trait, tol=1e-5, n_random_vectors=30)
    want = oracle.fit_ai_reml(oracle.Grm(d, af, oracle.GRM_F64), y, c.x, trait=trait, tol=1e-5)
    _say(f"tau gpu={got['tau']} oracle={want['tau']} it={got['iterations']}/{want['iterations']}")
    assert got["iterations"] == want["iterations"]
    assert got["converged"] == want["converged"]
    assert rel(got["tau"], want["tau"]).max() < 1e-6
    np.testing.assert_allclose(got["mu"], want["mu"], rtol=1e-6)
    np.testing.assert_allclose(got["alpha"], want["alpha"], rtol=1e-6, atol=1e-9)
    assert abs(got["log_likelihood"] - want["log_likelihood"]) < 1e-6 * abs(want["log_likelihood"])


# --------------------------------------------------------------------------- fit-null flow
def test_fit_null_vs_oracle():
    c = synth.unrelated(n=600, m_grm=2000, m_test=0, seed=11, h2=0.5, prevalence=0.2)
    idx = np.arange(c.n, dtype=np.uint32)
    got = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx)
    want = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, c.y, c.x)
    _say(f"fit_null gpu tau={got['tau']} vr={got['vr']} oracle tau={want['tau']} vr={want['vr']}")
    np.testing.assert_array_equal(got["marker_class"], want["marker_class"])
    assert (got["n_grm_markers"], got["n_vr_markers"]) == (want["n_grm_markers"], want["n_vr_markers"])
    assert got["iterations"] == want["iterations"]
    assert rel(got["tau"], want["tau"]).max() < 1e-6
    assert got["vr_markers_used"] == want["vr_markers_used"]
    assert abs(got["vr"] - want["vr"]) < 1e-6 * want["vr"]
    np.testing.assert_allclose(got["mu"], want["mu"], rtol=1e-6)
    np.testing.assert_allclose(got["xvx_inv_xv"], want["xvx_inv_xv"], rtol=1e-5, atol=1e-12)


# --------------------------------------------------------------------------- step 2
def _null(n, p, seed, eta=-2.5):
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(p - 1)])
    mu = 1 / (1 + np.exp(-(eta + 0.3 * X[:, -1] + 0.4 * rng.standard_normal(n))))
    y = (rng.random(n) < mu).astype(float)
    a = oracle.xvx_inv_xv(X.T, mu * (1 - mu))
    return X, mu, y, a


def _compare(got, want, mask_p=1e-10, integer_dosages=True, ac_rtol=1e-12):
    # counts are always exact; allele sums are exact for hard calls and sums of
    # non-integer dosages agree to rounding (a different summation order; the oracle adds the
    # n rounded q/127 in sequence, an error that grows with n)
    for f in ("n", "n_cases", "n_controls"):
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    for f in ("ac", "af", "mac", "af_cases", "af_controls"):
        if integer_dosages:
            np.testing.assert_array_equal(got[f], want[f], err_msg=f)
        else:
            np.testing.assert_allclose(got[f], want[f], rtol=ac_rtol, atol=1e-12, err_msg=f)
    for f in ("score", "var_t", "var_t_star", "beta"):
        np.testing.assert_allclose(got[f], want[f], rtol=1e-8, atol=1e-12, err_msg=f)
    np.testing.assert_array_equal(got["is_spa"], want["is_spa"])
    np.testing.assert_array_equal(got["spa_converged"], want["spa_converged"])
    m = want["pvalue"] >= mask_p
    err = np.abs(neglog10(got["pvalue"][m]) - neglog10(want["pvalue"][m]))
    assert err.max() < 1e-6, err.max()
    m2 = want["pvalue_noadj"] >= mask_p
    err2 = np.abs(neglog10(got["pvalue_noadj"][m2]) - neglog10(want["pvalue_noadj"][m2]))
    assert err2.max() < 1e-6


@pytest.mark.parametrize("p", [1, 3, 5])
def test_score_u8_vs_oracle(p):
    n = 3000
    X, mu, y, a = _null(n, p, seed=p)
    q, _ = synth.dosage_u8(n, 0, 300, seed=5, maf=(2e-3, 0.5))
    # plant strong signals so that SPA is exercised on both tails
    rng = np.random.default_rng(0)
    for v in range(0, 300, 7):
        q[v, (y > 0.5) & (rng.random(n) < 0.3)] = 254
    vr = 1.05
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X.T, a, vr, y=y, spa_pval_cutoff=0.05)
    got = e.test_markers_u8(q, 1.0 / 127.0)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.3), X.T, a, vr, y=y, spa_cutoff=0.05)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    _say(f"p={p} spa'd {int(want['is_spa'].sum())} of {len(q)}; min p {want['pvalue'].min():.3g}")
    assert want["is_spa"].sum() > 10
    _compare(got, want, integer_dosages=False)


def _score_u8_padded(e, q, scale=1.0 / 127.0):
    """Score (nvar, n) uint8 dosages through the device entry point with each row padded to a
    multiple of 16 bytes: the byte kernels of the sums and the SPA passes take 16-byte-aligned
    rows only (others take the generic loader kernels)."""
    from csaige_amd import hip
    nvar, n = q.shape
    ld = -(-n // 16) * 16
    qp = np.zeros((nvar, ld), np.uint8)
    qp[:, :n] = q
    rs = C.sizeof(sg._lib.VariantResult)
    d_q, d_r = hip.DeviceBuffer(qp.nbytes), hip.DeviceBuffer(nvar * rs)
    d_q.upload(qp)
    e.test_markers_u8_device(d_q.value, ld, nvar, scale, d_r.value)
    hip.device_synchronize()
    host = np.empty(nvar * rs, np.uint8)
    d_r.download(host)
    return sg.results_to_dict((sg._lib.VariantResult * nvar).from_buffer(host))


@pytest.mark.parametrize("p,form", [(1, 0), (1, 1), (1, 2), (3, 0), (3, 1), (3, 2), (8, 1),
                                    (8, 2)])
def test_score_u8_each_sums_form_vs_oracle(p, form):
    """Each kernel of the uint8 sums (0 lanes over samples, at most 3 covariates; 1 a lane per
    variant; 2 the int8 matrix cores on a fixed-point expansion of the features, 11 column
    tiles at the largest covariate count, 8) against the oracle, on a sample count that
    leaves a partial 256-sample block (rows padded to 16 bytes, which these kernels need),
    hard calls and noisy dosages; the generic loader kernel (unpadded rows) beside them."""
    n, nvar = 5000 + 77, 700
    X, mu, y, a = _null(n, p, seed=60 + p)
    q, _ = synth.dosage_u8(n, 0, nvar, seed=21 + p, maf=(1e-3, 0.5))
    rng = np.random.default_rng(p)
    for v in range(0, nvar, 9):
        q[v, (y > 0.5) & (rng.random(n) < 0.3)] = 254
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.05, y=y, spa_pval_cutoff=0.05)
    e.set_sums_form(form)
    got = _score_u8_padded(e, q)
    generic = e.test_markers_u8(q, 1.0 / 127.0)  # n % 16 != 0: unaligned rows
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.05, y=y, spa_cutoff=0.05)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    _say(f"p={p} form {form}: spa'd {int(want['is_spa'].sum())} of {nvar}")
    assert want["is_spa"].sum() > 10
    _compare(got, want, integer_dosages=False)
    _compare(generic, want, integer_dosages=False)


def test_score_u8_timed_shape_matches_fixed_forms():
    """At 70,000 samples the first batch times every form and split count (each computes the
    full sums) and keeps the fastest; its results equal the fixed forms' to rounding, and the
    allele counts of the int8 matrix-core form are the exact byte sums (scaled per split)."""
    n, nvar = 70000, 512
    X, mu, y, a = _null(n, 1, seed=77)
    q, _ = synth.dosage_u8(n, 0, nvar, seed=78, maf=(1e-4, 0.5))
    res = {}
    for form in (-1, 0, 2):
        e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.0, y=y, spa_pval_cutoff=0.05)
        e.set_sums_form(form)
        res[form] = e.test_markers_u8(q, 1.0 / 127.0)
        res[form] = e.test_markers_u8(q, 1.0 / 127.0)  # the batch after the shape search
    ac = q.astype(np.int64).sum(axis=1) / 127.0
    np.testing.assert_allclose(res[2]["ac"], ac, rtol=1e-14, atol=0)
    for form in (-1, 0):
        _compare(res[form], res[2], integer_dosages=False)


def test_step2_full_sample_count_vs_oracle():
    """Step 2 at the streamed workload's own size: 400,000 samples, an intercept-only null of
    about 1% cases, uint8 dosages from the same generator and seeds as the benchmark, SPA
    where |z| > 2.  The first batch times the sums kernels, the second is scored with the
    kept shape (sample splits of a full-size batch, the zero-byte table, carrier-only SPA
    passes); both against the oracle."""
    n, nvar = 400000, 512
    nf = synth.null_fit(n=n, p=1, eta_mean=-4.6, eta_sd=0.5, seed=3)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    q, _ = synth.dosage_u8(n, 0, nvar, seed=4)
    cutoff = 0.04550026389635842  # 2 (1 - Phi(2))
    e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                           spa_pval_cutoff=cutoff)
    assert e.set_spa_split(True)
    first = e.test_markers_u8(q, 1.0 / 127.0)
    got = e.test_markers_u8(q, 1.0 / 127.0)
    eo = oracle.ScoreEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                            spa_cutoff=cutoff)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    _say(f"n=400,000: {int(want['is_spa'].sum())} of {nvar} SPA'd")
    assert want["is_spa"].sum() >= 5
    # sequential sums of 400,000 rounded terms in the oracle: up to n 2^-53 = 4.4e-11 apart
    for r in (first, got):
        _compare(r, want, integer_dosages=False, ac_rtol=5e-11)


def test_step2_full_sample_count_with_covariates_vs_oracle():
    """Step 2 with two covariates beside the intercept at 400,000 samples: the sums kernels
    of p = 3 (seven linear features, 72 digit columns on the matrix cores) at the real
    sample-split bound, the full SPA passes of a covariate model, first (timed shapes) and
    second batch against the oracle."""
    n, nvar = 400000, 384
    X, mu, y, a = _null(n, 3, seed=41, eta=-4.0)
    q, _ = synth.dosage_u8(n, 0, nvar, seed=42, maf=(1e-3, 0.5))
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.4), X.T, a, 1.02, y=y, spa_pval_cutoff=0.05)
    first = e.test_markers_u8(q, 1.0 / 127.0)
    got = e.test_markers_u8(q, 1.0 / 127.0)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.4), X.T, a, 1.02, y=y, spa_cutoff=0.05)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    _say(f"n=400,000, p=3: {int(want['is_spa'].sum())} of {nvar} SPA'd")
    assert want["is_spa"].sum() >= 5
    for r in (first, got):
        _compare(r, want, integer_dosages=False, ac_rtol=5e-11)
    e.close()


# With the score near the null mean the saddle point sits near 0 and the Lugannani-Rice form
# z = w + ln(v / w) / w (fast.rs:135-180) turns a rounding error dK of the summed K(zeta) into
# dp ~ phi(w) dK / w^3: with dK ~ 1e-13 (sums over 10^3 to 10^5 samples) the p of |z| below
# about 0.005 is rounding noise of the summation order, on the CPU as on the GPU, so there
# only the flags are compared (tstat = z^2).
_LR_WELL_CONDITIONED = 2.5e-5


def test_spa_carrier_passes_vs_full_passes_and_oracle():
    """Intercept-only engines take the SPA passes of variants with at most 35% carriers over
    the carriers only, the dosage-0 samples entering through a table of sums over all
    samples (Chebyshev series of the exact sums).  At 60,000 samples with every variant
    SPA'd (hard calls from rare to common, noisy dosages, planted signals on both tails) the
    results equal those of the passes over every sample to rounding, and the oracle's."""
    n, nvar = 60000, 320
    X, mu, y, a = _null(n, 1, seed=91)
    q, _ = synth.dosage_u8(n, 0, nvar, seed=92, maf=(2e-4, 0.5))
    rng = np.random.default_rng(93)
    for v in range(0, nvar, 5):
        q[v, (y > 0.5) & (rng.random(n) < 0.2)] = 254
    for v in range(2, nvar, 11):
        q[v, (y < 0.5) & (rng.random(n) < 0.01)] = 254
    res = {}
    for split in (True, False):
        e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y, spa_pval_cutoff=1.0)
        assert e.set_spa_split(split)
        res[split] = e.test_markers_u8(q, 1.0 / 127.0)
    got, full = res[True], res[False]
    carriers = (q > 0).mean(axis=1)
    _say(f"variants with <= 35% carriers: {int((carriers <= 0.35).sum())} of {nvar}; "
         f"failed roots {int((full['spa_converged'] == 0).sum())}")
    assert (carriers <= 0.35).sum() > nvar // 2
    np.testing.assert_array_equal(got["is_spa"], full["is_spa"])
    np.testing.assert_array_equal(got["spa_converged"], full["spa_converged"])
    cond = full["tstat"] > _LR_WELL_CONDITIONED
    d = np.abs(neglog10(got["pvalue"]) - neglog10(full["pvalue"]))
    strong = full["tstat"] > 1.0
    _say(f"max |d -log10 p| carrier vs full passes: {d[cond].max():.3g} (|z| > 0.005), "
         f"{d[strong].max():.3g} (|z| > 1)")
    assert d[cond].max() < 1e-6 and d[strong].max() < 1e-10
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y, spa_cutoff=1.0)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    np.testing.assert_array_equal(got["spa_converged"], want["spa_converged"])
    _compare({k: v[cond] for k, v in got.items()}, {k: v[cond] for k, v in want.items()},
             integer_dosages=False)


@pytest.mark.parametrize("p", [1, 2])
def test_score_u8_every_variant_spa_vs_oracle(p):
    """SPA on every variant (cutoff 1): a batch of 2,500 gated variants, more than the 1,024
    variant-slices of a Newton pass, so the passes start with one slice per variant and
    split the variants still searching into more slices as the others finish; intercept-only
    (byte-table passes) and with a covariate (per-sample passes)."""
    n, nvar = 3000, 2500
    X, mu, y, a = _null(n, p, seed=40 + p)
    q, _ = synth.dosage_u8(n, 0, nvar, seed=9 + p, maf=(2e-3, 0.5))
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y, spa_pval_cutoff=1.0)
    got = e.test_markers_u8(q, 1.0 / 127.0)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y, spa_cutoff=1.0)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    _say(f"p={p}: {int(want['is_spa'].sum())} of {nvar} SPA'd, "
         f"{int((want['spa_converged'] == 0).sum())} failed roots")
    assert want["is_spa"].sum() > 1024
    cond = want["tstat"] > _LR_WELL_CONDITIONED
    _say(f"ill-conditioned saddle points (|z| < 0.005): {int((~cond).sum())}")
    assert cond.sum() > nvar - 20
    np.testing.assert_array_equal(got["is_spa"], want["is_spa"])
    np.testing.assert_array_equal(got["spa_converged"], want["spa_converged"])
    _compare({k: v[cond] for k, v in got.items()}, {k: v[cond] for k, v in want.items()},
             integer_dosages=False)


@pytest.mark.parametrize("prev,p", [(0.03, 1), (0.01, 1), (0.03, 2)])
def test_score_u8_spa_failed_roots_vs_oracle(prev, p):
    """Rare variants under extreme case imbalance, where the safeguarded Newton of many SPA
    roots fails (its iterates run off to where e^{-g t} overflows): flags, counts and
    p-values against the oracle, for the byte-table passes (intercept only) and the
    per-sample passes (a covariate)."""
    n = 4000
    eta = math.log(prev / (1 - prev))
    nf = synth.null_fit(n=n, p=p, eta_mean=eta, eta_sd=0.5, seed=3)
    a = oracle.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    q = np.zeros((2000, n), np.uint8)
    synth.dosage_u8(n, 0, 2000, seed=4, maf=(1e-4, 0.5), out=q)
    e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                           spa_pval_cutoff=0.05)
    got = e.test_markers_u8(q, 1.0 / 127.0)
    eo = oracle.ScoreEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                            spa_cutoff=0.05)
    want = eo.test_batch_u8(q, 1.0 / 127.0)
    failed = int((want["spa_converged"] == 0).sum())
    _say(f"prev={prev} p={p}: {int(want['is_spa'].sum())} SPA'd, {failed} failed roots")
    assert failed >= 10
    _compare(got, want, integer_dosages=False)


def test_score_f64_missing_and_categorical_vr():
    n = 700
    X, mu, y, a = _null(n, 2, seed=9)
    rng = np.random.default_rng(1)
    G = rng.binomial(2, rng.uniform(0.001, 0.3, (40, 1)), (40, n)).astype(float)
    G[5, 10] = np.nan
    G[6] = 0.0
    G[7, y > 0.5] = 2.0
    cat = [(1.5, 0.9), (10.5, 1.2), (math.inf, 1.0)]
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.0, y=y, categorical_vr=cat)
    got = e.test_markers(G)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.0, y=y, cat_vr=cat)
    want_rows = [eo.test_marker(g) for g in G]
    want = {k: np.array([r[k] for r in want_rows]) for k in want_rows[0]}
    assert np.isnan(got["score"][5]) and got["pvalue"][5] == 1.0
    got["score"][5] = want["score"][5] = 0.0
    got["var_t"][5] = want["var_t"][5] = 0.0
    got["var_t_star"][5] = want["var_t_star"][5] = 0.0
    _compare(got, want)


def test_score_plink_hard_calls():
    n_total = 900
    body, bpm = random_bed(n_total, 120, seed=4)
    idx = np.arange(n_total - 1, -1, -1, dtype=np.uint32)[:850].copy()
    X, mu, y, a = _null(len(idx), 3, seed=8)
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.1, y=y)
    got = e.test_markers_plink(body, bpm, n_total, 120, sample_idx=idx)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.3), X.T, a, 1.1, y=y)
    rows = [eo.test_marker(oracle.decode_marker(body[j], n_total, idx)) for j in range(120)]
    want = {k: np.array([r[k] for r in rows]) for k in rows[0]}
    _compare(got, want)


def test_spa_single_vs_oracle():
    rng = np.random.default_rng(6)
    n = 2000
    mu = rng.uniform(0.005, 0.1, n)
    g = rng.binomial(2, 0.01, n) - 0.02 * rng.random(n)
    m1 = mu @ g
    sd = math.sqrt(np.sum(mu * (1 - mu) * g * g))
    for z in (2.5, 4.0, 8.0, -3.0):
        q = m1 + z * sd
        for fast in (True, False):
            a = sg.spa_binary_fast(mu, g, q, 2 * m1 - q, 0.01, fast=fast)
            b = oracle.spa_binary(mu, g, q, 2 * m1 - q, 0.01, fast=fast)
            assert (a["converged"], a["is_spa"], a["iters1"], a["iters2"]) == \
                   (b["converged"], b["is_spa"], b["iters1"], b["iters2"])
            assert abs(neglog10(a["pvalue"]) - neglog10(b["pvalue"])) < 1e-6
    # q outside the support: root at infinity, both tails fall back to p_noadj / 2
    a = sg.spa_binary_fast(np.full(10, 0.5), np.ones(10), 20.0, -15.0, 0.01)
    assert a["converged"] and math.isinf(a["root1"]) and a["pvalue"] == 0.01


def test_pipeline_fit_null_then_score():
    """fit-null on the GPU then step 2 on the GPU vs the same two steps of the oracle."""
    c = synth.unrelated(n=800, m_grm=2000, m_test=300, seed=21, h2=0.4, prevalence=0.1)
    idx = np.arange(c.n, dtype=np.uint32)
    fg = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx)
    fo = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, c.y, c.x)
    tau = fg["tau"]
    e = sg.ScoreTestEngine(fg["mu"], c.y - fg["mu"], tau, c.x, fg["xvx_inv_xv"], fg["vr"], y=c.y)
    got = e.test_markers_plink(c.bed_test, c.bpm, c.n, c.bed_test.shape[0])
    eo = oracle.ScoreEngine(fo["mu"], c.y - fo["mu"], fo["tau"], c.x, fo["xvx_inv_xv"], fo["vr"], y=c.y)
    rows = [eo.test_marker(oracle.decode_marker(b, c.n)) for b in c.bed_test]
    want = {k: np.array([r[k] for r in rows]) for k in rows[0]}
    m = want["pvalue"] >= 1e-10
    err = np.abs(neglog10(got["pvalue"][m]) - neglog10(want["pvalue"][m]))
    _say(f"pipeline max |d -log10 p| = {err.max():.3g}")
    assert err.max() < 1e-5


def test_model_file_step1_to_step2(tmp_path):
    """fit-null on the GPU written as a .saige.model (with alpha and the per-marker variance
    ratios against the oracle), read back, and step 2 from the loaded model equal to step 2
    from the fit itself; the SAIGE result lines carry the same numbers."""
    from csaige_amd import model as M
    c = synth.unrelated(n=800, m_grm=2000, m_test=120, seed=21, h2=0.4, prevalence=0.1)
    idx = np.arange(c.n, dtype=np.uint32)
    fg = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx)
    fo = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, c.y, c.x)
    np.testing.assert_allclose(fg["alpha"], fo["alpha"], rtol=1e-6, atol=1e-9)
    assert len(fg["per_marker_vr"]) == fg["vr_markers_used"] == len(fo["per_marker_vr"])
    np.testing.assert_allclose(fg["per_marker_vr"], fo["per_marker_vr"], rtol=1e-6)
    ids = [f"sample{i}" for i in range(c.n)]
    m = M.NullModel.from_fit_null(fg, c.y, c.x, ids)
    M.save_model(m, tmp_path / "gpu.saige.model")
    m2 = M.load_model(tmp_path / "gpu.saige.model")
    assert m2.sample_ids == ids and m2.tau == tuple(fg["tau"])
    assert abs(m2.variance_ratio.variance_ratio - fg["vr"]) == 0.0
    e1 = sg.ScoreTestEngine(fg["mu"], c.y - fg["mu"], fg["tau"], c.x, fg["xvx_inv_xv"], fg["vr"],
                            y=c.y)
    e2 = M.engine_from_model(m2)
    got1 = e1.test_markers_plink(c.bed_test, c.bpm, c.n, c.bed_test.shape[0])
    got2 = e2.test_markers_plink(c.bed_test, c.bpm, c.n, c.bed_test.shape[0])
    for k in ("ac", "score", "var_t", "pvalue", "pvalue_noadj", "is_spa", "beta", "se_beta"):
        np.testing.assert_array_equal(got1[k], got2[k], err_msg=k)
    nv = len(got2["pvalue"])
    M.write_results(tmp_path / "out.txt", got2, ["1"] * nv, np.arange(nv) + 1,
                    [f"rs{v}" for v in range(nv)], ["A"] * nv, ["G"] * nv)
    lines = (tmp_path / "out.txt").read_text().splitlines()
    assert lines[0].split() == M.HEADER.split() and len(lines) == nv + 1
    cols = M.HEADER.split()
    back = np.array([float(ln.split()[cols.index("p.value")]) for ln in lines[1:]])
    np.testing.assert_array_equal(back, got2["pvalue"])  # shortest round-trip digits


@pytest.mark.parametrize("prec,mode", [(sg.PREC_FP64, oracle.GRM_F64),
                                       (sg.PREC_F32, oracle.GRM_F32)], ids=["fp64", "f32"])
def test_unrelated_2000_workload_end_to_end(prec, mode):
    """The 2,000-sample CPU-runnable workload in full: 2,000 unrelated samples, 5,000 GRM
    markers, 2,000 test variants, 2 covariates, liability h2 = 0.3, top 10% cases; fit-null
    (AI-REML with PCG tol 1e-5 and 30 probes) then score test + SPA of every test variant,
    GPU against the oracle: tau and VR to 1e-6 relative, -log10 p to 1e-6, counts exact.
    fp64: the GRM in fp64 against the oracle's fp64 GRM; f32: saige-rs's own f32 GRM
    arithmetic against the oracle's restatement of it (the reference's results)."""
    c = synth.unrelated(n=2000, m_grm=5000, m_test=2000, seed=1, h2=0.3, prevalence=0.10)
    idx = np.arange(c.n, dtype=np.uint32)
    fg = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx, precision=prec)
    fo = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, c.y, c.x, grm_mode=mode)
    _say(f"2000-sample fit ({prec}): gpu tau={fg['tau']} vr={fg['vr']} it={fg['iterations']}; "
         f"oracle tau={fo['tau']} vr={fo['vr']} it={fo['iterations']}")
    np.testing.assert_array_equal(fg["marker_class"], fo["marker_class"])
    assert fg["iterations"] == fo["iterations"]
    assert rel(fg["tau"], fo["tau"]).max() < 1e-6
    assert abs(fg["vr"] - fo["vr"]) < 1e-6 * fo["vr"]
    e = sg.ScoreTestEngine(fg["mu"], c.y - fg["mu"], fg["tau"], c.x, fg["xvx_inv_xv"], fg["vr"],
                           y=c.y)
    got = e.test_markers_plink(c.bed_test, c.bpm, c.n, c.bed_test.shape[0])
    eo = oracle.ScoreEngine(fo["mu"], c.y - fo["mu"], fo["tau"], c.x, fo["xvx_inv_xv"], fo["vr"],
                            y=c.y)
    rows = [eo.test_marker(oracle.decode_marker(b, c.n)) for b in c.bed_test]
    want = {k: np.array([r[k] for r in rows]) for k in rows[0]}
    np.testing.assert_array_equal(got["ac"], want["ac"])
    np.testing.assert_array_equal(got["mac"], want["mac"])
    np.testing.assert_array_equal(got["is_spa"], want["is_spa"])
    m = want["pvalue"] >= 1e-10
    err = np.abs(neglog10(got["pvalue"][m]) - neglog10(want["pvalue"][m]))
    _say(f"2000-sample step 2: {int(want['is_spa'].sum())} SPA'd, max |d -log10 p| = {err.max():.3g}")
    assert err.max() < 1e-6


# --------------------------------------------------------------------------- mixed precision
# Stated error bounds of the reduced-precision products against the fp64 oracle, relative
# to the column norm: bf16 hi+lo with fp32 accumulation 2e-5; three int8 slices with exact
# int32 accumulation 1e-6 (per-entry operand error <= 6.1e-8 of the column's largest entry);
# two int8 slices 1e-4 (per-entry operand error <= 1.55e-5 of the column's largest entry, in
# each of the two operand passes).
_BOUND = {sg.PREC_MIXED: 2e-5, sg.PREC_INT8: 1e-6, sg.PREC_INT8X2: 1e-4}
_REDUCED = [sg.PREC_MIXED, sg.PREC_INT8, sg.PREC_INT8X2]
_REDUCED_IDS = ["bf16", "int8", "int8x2"]


@pytest.mark.parametrize("prec", _REDUCED, ids=_REDUCED_IDS)
@pytest.mark.parametrize("n,m,k,p_missing", [(1000, 1300, 1, 0.0), (2048, 3000, 34, 0.0),
                                             (700, 513, 100, 0.0), (1100, 1500, 136, 0.0),
                                             (900, 1400, 34, 0.03), (1300, 700, 7, 0.15)],
                         ids=["k1", "k34", "k100", "k136", "k34-missing3pct",
                              "k7-missing15pct"])
def test_grm_mixed_error_bound(n, m, k, p_missing, prec):
    """Exact genotypes on MFMA with the standardization factored out: stated error against
    the fp64 oracle.  k = 100 and 136 run the int8 paths as one launch of several column
    groups (int8: 300 and 408 slice columns; int8x2: 200 and 272).  With missing calls the
    sparse per-marker and per-sample corrections restore the reference's mean imputation
    (pcg.rs:108-113), up to 15% missing (the fit-null QC limit, fit_null.rs:209-215)."""
    _say(f"mixed grm prec={prec} n={n} m={m} k={k} missing={p_missing}")
    body, bpm, d, af = _grm_pair(n, m, seed=3 * n + m, p_missing=p_missing, mono=True)
    g = sg.OnTheFlyGrm(body, bpm, n, af, precision=prec)
    ref = oracle.Grm(d, af, oracle.GRM_F64)
    rng = np.random.default_rng(k)
    V = rng.standard_normal((k, n)) + 3.0  # a common offset exercises the centring
    got = g.mat_mat(V)
    worst = 0.0
    for c in range(k):
        want = ref.mat_vec(V[c])
        worst = max(worst, np.linalg.norm(got[c] - want) / np.linalg.norm(want))
    _say(f"prec {prec} grm worst relative error {worst:.2e}")
    assert worst < _BOUND[prec]
    # one column alone gives the same values as inside the block (int8: bit for bit,
    # the integer partial sums do not depend on how the work is split)
    one = g.mat_mat(V[k // 2:k // 2 + 1])[0]
    if prec in (sg.PREC_INT8, sg.PREC_INT8X2):
        np.testing.assert_array_equal(one, got[k // 2])
    else:
        assert np.linalg.norm(one - got[k // 2]) < 1e-5 * np.linalg.norm(one)


@pytest.mark.parametrize("k", [1, 6, 15, 20, 25, 30, 35, 40, 45, 50, 60])
def test_grm_int8_every_tile_width(k):
    """The int8 GEMM is instantiated per column-tile count KT = 1..10 of a group (three
    slices: 3k columns in ceil(3k / 16) tiles, one group up to 10 tiles; k = 60 is two groups
    of 6 tiles).  Every width meets the three-slice bound against the fp64 oracle, and a
    column computed inside the wide launch equals, bit for bit, the same column computed
    alone (KT = 1): the int32 sums are exact whatever the tile shape, the split of the
    reduction or the group."""
    n, m = 900, 1100
    _say(f"int8 tile width k={k} ({-(-3 * k // 16)} tiles)")
    body, bpm, d, af = _grm_pair(n, m, seed=11 * k + 5, p_missing=0.02, mono=True)
    g = sg.OnTheFlyGrm(body, bpm, n, af, precision=sg.PREC_INT8)
    ref = oracle.Grm(d, af, oracle.GRM_F64)
    rng = np.random.default_rng(100 + k)
    V = rng.standard_normal((k, n)) + 1.5
    got = g.mat_mat(V)
    worst = 0.0
    for c in range(k):
        want = ref.mat_vec(V[c])
        worst = max(worst, np.linalg.norm(got[c] - want) / np.linalg.norm(want))
    _say(f"int8 k={k} worst relative error {worst:.2e}")
    assert worst < _BOUND[sg.PREC_INT8]
    for c in sorted({0, k // 2, k - 1}):
        np.testing.assert_array_equal(g.mat_mat(V[c:c + 1])[0], got[c])


@pytest.mark.parametrize("prec", _REDUCED, ids=_REDUCED_IDS)
@pytest.mark.parametrize("p_missing", [0.0, 0.03], ids=["complete", "missing3pct"])
def test_fit_null_mixed_vs_fp64_oracle(prec, p_missing):
    """Mixed-precision fit-null against the FP64 oracle: tau and VR within 1e-4, AI-REML
    iterations within one of the oracle's.  With 3% missing calls among the GRM markers the
    reduced forms run with the sparse missing-call corrections.  (Missing calls go to GRM
    markers only: saige-rs passes a VR marker's NaN dosages on into estimate_variance_ratio,
    whose NaN ratio then fails every CV test, so that every VR candidate is solved to the
    500-iteration PCG limit and the ratio is NaN (variance_ratio.rs:130-211); the oracle
    reproduces that, at minutes of CPU time.)"""
    c = synth.unrelated(n=1500, m_grm=3000, m_test=0, seed=12, h2=0.5, prevalence=0.2)
    bed = c.bed.copy()
    if p_missing:
        base = _oracle_fit(("unrelated1500", 0.0), c.bed, c.bpm, c.n, c.y, c.x)
        af = sg.marker_stats(c.bed, c.bpm, c.n, c.bed.shape[0])[2]
        # GRM markers far from the MAF filter keep their class with the calls removed
        rows = np.flatnonzero((base["marker_class"] == 1) & (af > 0.05) & (af < 0.95))
        codes = np.stack([(bed >> (2 * t)) & 3 for t in range(4)], axis=2)
        rng = np.random.default_rng(77)
        hit = rng.random(codes[rows].shape) < p_missing
        sub = codes[rows]
        sub[hit] = 1
        codes[rows] = sub
        bed = (codes[..., 0] | (codes[..., 1] << 2) | (codes[..., 2] << 4)
               | (codes[..., 3] << 6)).astype(np.uint8)
    idx = np.arange(c.n, dtype=np.uint32)
    got = sg.fit_null(bed, c.bpm, c.n, bed.shape[0], c.y, c.x, sample_idx=idx, precision=prec)
    want = _oracle_fit(("unrelated1500", p_missing), bed, c.bpm, c.n, c.y, c.x)
    _say(f"prec {prec} missing {p_missing} tau={got['tau'][1]:.8g} oracle={want['tau'][1]:.8g} "
         f"vr {got['vr']:.8g}/{want['vr']:.8g} it {got['iterations']}/{want['iterations']}")
    np.testing.assert_array_equal(got["marker_class"], want["marker_class"])
    assert got["converged"]
    assert rel(got["tau"], want["tau"]).max() < 1e-4
    assert abs(got["vr"] - want["vr"]) < 1e-4 * want["vr"]
    assert abs(got["iterations"] - want["iterations"]) <= 1


@pytest.mark.parametrize("prec", [sg.PREC_INT8, sg.PREC_INT8X2], ids=["int8", "int8x2"])
def test_fit_null_int8_small_cohort_vs_fp64_oracle(prec):
    """The cohort on which plain two-slice solves drifted (n = 1,600, tau about 0.16): the
    int8x2 fit takes three operand slices for the narrow solves and, once the relative tau
    change is below 1e-3, for the y~ and X columns of the main block solves (the probe
    columns keep two), and meets tau within 1e-4 of the FP64 oracle with the oracle's AI-REML
    iteration count within one, for every trait."""
    c = synth.biobank(n=1600, m=3000, frac_related=0.15, prevalences=(0.05, 0.1, 0.3), seed=21,
                      h2=0.5, n_causal=500)
    idx = np.arange(c.n, dtype=np.uint32)
    multi = sg.fit_null_multi(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                              precision=prec)
    for t in range(c.y.shape[0]):
        want = _oracle_fit(("biobank1600", t), c.bed, c.bpm, c.n, c.y[t], c.x)
        got = multi[t]
        _say(f"prec {prec} trait {t}: tau {got['tau'][1]:.8g} oracle {want['tau'][1]:.8g} "
             f"rel {rel(got['tau'], want['tau']).max():.2e} it {got['iterations']}/"
             f"{want['iterations']} vr rel {abs(got['vr'] - want['vr']) / want['vr']:.2e}")
        assert got["converged"]
        assert rel(got["tau"], want["tau"]).max() < 1e-4
        assert abs(got["vr"] - want["vr"]) < 1e-4 * want["vr"]
        assert abs(got["iterations"] - want["iterations"]) <= 1


# --------------------------------------------------------------------------- several phenotypes
@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_INT8, sg.PREC_INT8X2],
                         ids=["fp64", "int8", "int8x2"])
def test_fit_null_multi_matches_single_trait(prec):
    """Three phenotypes fitted together (joined GRM products) follow their single-trait
    fits: same AI-REML and PCG iteration counts, same tau / VR / mu to 1e-10."""
    c = synth.biobank(n=1600, m=3000, frac_related=0.15, prevalences=(0.05, 0.1, 0.3), seed=21,
                      h2=0.5, n_causal=500)
    idx = np.arange(c.n, dtype=np.uint32)
    rng = np.random.default_rng(4)
    Y = c.y.copy()
    Y[2] = Y[2] + 0.5 * rng.standard_normal(c.n)  # a quantitative trait among binary ones
    traits = [sg.TRAIT_BINARY, sg.TRAIT_BINARY, sg.TRAIT_QUANTITATIVE]
    multi = sg.fit_null_multi(c.bed, c.bpm, c.n, c.bed.shape[0], Y, c.x, traits=traits,
                              sample_idx=idx, precision=prec)
    for t in range(3):
        one = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], Y[t], c.x, sample_idx=idx,
                          trait=traits[t], precision=prec)
        _say(f"prec {prec} trait {t}: tau multi={multi[t]['tau']} single={one['tau']} "
             f"it {multi[t]['iterations']}/{one['iterations']} pcg {multi[t]['pcg_iterations']}"
             f"/{one['pcg_iterations']}")
        assert multi[t]["converged"] and one["converged"]
        assert multi[t]["iterations"] == one["iterations"]
        assert multi[t]["pcg_iterations"] == one["pcg_iterations"]
        assert multi[t]["grm_products"] == one["grm_products"]
        assert rel(multi[t]["tau"], one["tau"]).max() < 1e-10
        assert abs(multi[t]["vr"] - one["vr"]) < 1e-10 * one["vr"]
        np.testing.assert_allclose(multi[t]["mu"], one["mu"], rtol=1e-10)
        np.testing.assert_array_equal(multi[t]["marker_class"], one["marker_class"])
    if prec == sg.PREC_FP64:
        want = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, Y[0], c.x)
        assert rel(multi[0]["tau"], want["tau"]).max() < 1e-6
        assert abs(multi[0]["vr"] - want["vr"]) < 1e-6 * want["vr"]


# --------------------------------------------------------------------------- launch log
def test_launch_log_files_launches_by_operand_width():
    """The process-wide launch log times the GEMM launches of a product under their operand
    width (k = 34 with two int8 slices: 80 columns; two launches per product) and also sees
    the launches of the GRM handle that a fit creates internally; a width filter keeps only
    the launches of that width."""
    import ctypes as C
    L = sg.load()
    body, bpm, d, af = _grm_pair(1200, 700, seed=5, p_missing=0.0)
    g = sg.OnTheFlyGrm(body, bpm, 1200, af, precision=sg.PREC_INT8X2)
    V = np.random.default_rng(2).standard_normal((34, 1200))
    g.mat_mat(V)  # warm-up outside the log
    ms, cnt = C.c_double(), C.c_uint64()
    sg.check(L.saige_launch_log_enable(1, 0))
    try:
        g.mat_mat(V)
        g.mat_mat(V[:3])  # 6 slice columns: width 16
        sg.check(L.saige_launch_log_read(80, C.byref(ms), C.byref(cnt)))
        assert cnt.value == 2 and ms.value > 0.0
        sg.check(L.saige_launch_log_read(0, C.byref(ms), C.byref(cnt)))
        assert cnt.value == 0  # the read cleared the log, including the other width
        c = synth.unrelated(n=600, m_grm=800, m_test=0, seed=3, h2=0.5, prevalence=0.2)
        idx = np.arange(c.n, dtype=np.uint32)
        r = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                        precision=sg.PREC_INT8X2)
        sg.check(L.saige_launch_log_read(0, C.byref(ms), C.byref(cnt)))
        _say(f"launch log: fit-null {cnt.value} GEMM launches, {ms.value:.2f} ms, "
             f"{r['grm_products']} column-products")
        assert cnt.value >= 2 and ms.value > 0.0
        sg.check(L.saige_launch_log_enable(1, 16))  # narrow launches only
        g.mat_mat(V)
        g.mat_mat(V[:3])
        sg.check(L.saige_launch_log_read(0, C.byref(ms), C.byref(cnt)))
        assert cnt.value == 2
    finally:
        sg.check(L.saige_launch_log_enable(0, 0))
        g.close()


def test_launch_log_summary_per_width():
    """The per-width summary files each GEMM launch under its operand width with the number of
    vectors it carried: k = 34 with two int8 slices is one launch pair at 80 columns carrying
    34 vectors each; 136 vectors run as one launch pair of three column groups (288 columns);
    the summary drains the log."""
    body, bpm, d, af = _grm_pair(1200, 700, seed=5, p_missing=0.0)
    g = sg.OnTheFlyGrm(body, bpm, 1200, af, precision=sg.PREC_INT8X2)
    V = np.random.default_rng(2).standard_normal((136, 1200))
    g.mat_mat(V)  # warm-up outside the log
    sg.launch_log_enable(True, 0)
    try:
        g.mat_mat(V[:34])
        g.mat_mat(V)
        s = sg.launch_log_summary()
        _say(f"launch log summary: {s}")
        assert set(s) == {80, 288}
        assert s[80][1] == 2 and s[80][2] == 68
        assert s[288][1] == 2 and s[288][2] == 272
        assert all(v[0] > 0.0 for v in s.values())
        assert sg.launch_log_summary() == {}
    finally:
        sg.launch_log_enable(False)
        g.close()


# --------------------------------------------------------------------------- PCG graphs
@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_INT8X2], ids=["fp64", "int8x2"])
def test_pcg_graphs_match_uncaptured_launches(prec):
    """PCG iterations captured in HIP graphs and replayed give the uncaptured launches'
    results bit for bit: repeated solves on one handle (the first eager, the second captured,
    the later ones replayed, with eager work and host reads between them), and whole fit-null
    runs with the graphs on, off and on again."""
    n, m = 1100, 1500
    body, bpm, d, af = _grm_pair(n, m, seed=31, p_missing=0.0)
    g = sg.OnTheFlyGrm(body, bpm, n, af, precision=prec)
    rng = np.random.default_rng(3)
    w = rng.uniform(0.05, 0.25, n)
    B = rng.standard_normal((34, n))
    prev = sg.pcg_graphs(False)
    try:
        want = sg.solve_sigma(g, w, (1.0, 0.4), B)
        sg.pcg_graphs(True)
        for rep in range(4):
            got = sg.solve_sigma(g, w, (1.0, 0.4), B)
            np.testing.assert_array_equal(got[0], want[0], err_msg=f"solve {rep}")
            np.testing.assert_array_equal(got[1], want[1])
            g.mat_vec(B[0])  # eager work between the replays
        c = synth.unrelated(n=1500, m_grm=3000, m_test=0, seed=12, h2=0.5, prevalence=0.2)
        idx = np.arange(c.n, dtype=np.uint32)
        runs = []
        for on in (True, False, True):
            sg.pcg_graphs(on)
            runs.append(sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                                    precision=prec))
        for r in (runs[0], runs[2]):
            _say(f"graphs prec {prec}: tau {r['tau'][1]!r} vs {runs[1]['tau'][1]!r}")
            for key in ("tau", "mu", "per_marker_vr"):
                np.testing.assert_array_equal(r[key], runs[1][key], err_msg=key)
            for key in ("vr", "iterations", "pcg_iterations", "grm_products"):
                assert r[key] == runs[1][key], key
    finally:
        sg.pcg_graphs(prev)
        g.close()


def test_fit_null_categorical_variance_ratio():
    """--use-categorical-vr (fit_null.rs:94-96, 348): the MAC-binned ratios
    (variance_ratio.rs:221-230, 263-287) of the markers the VR estimate used, in the order it
    used them (the ChaCha8 shuffle of the VR pool); the overall ratio is unchanged."""
    c = synth.unrelated(n=1500, m_grm=3000, m_test=0, seed=12, h2=0.5, prevalence=0.2)
    idx = np.arange(c.n, dtype=np.uint32)
    plain = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx)
    got = sg.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], c.y, c.x, sample_idx=idx,
                      use_categorical_vr=True)
    assert got["vr"] == plain["vr"] and plain["categorical_vr"] == []
    pool = np.flatnonzero(got["marker_class"] == 2)
    order = oracle.shuffle(12345, len(pool)).astype(np.int64)
    used = pool[order[:got["vr_markers_used"]]]
    mac = sg.marker_stats(c.bed, c.bpm, c.n, c.bed.shape[0])[3][used]
    want = oracle.categorical_vr(got["per_marker_vr"], mac)
    _say(f"categorical VR: {got['categorical_vr']} (want {want}) over {len(used)} markers")
    assert len(want) >= 1
    assert got["categorical_vr"] == want


def test_score_pool_equals_one_engine_and_oracle():
    """A pool of three engines (own streams, own host threads) over device-resident uint8
    batches of unequal sizes: every result equals, bit for bit, that of one engine holding the
    same tuned shapes, and the first batch matches the oracle."""
    from csaige_amd import hip
    n, sizes = 100000, (512, 512, 300, 512, 512, 512, 173, 300)
    nf = synth.null_fit(n=n, p=1, eta_mean=-4.6, eta_sd=0.5, seed=31)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    cutoff = 0.04550026389635842
    args = (nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05)
    V = sum(sizes)
    q, _ = synth.dosage_u8(n, 0, V, seed=32, maf=(1e-3, 0.5))
    d_q = hip.DeviceBuffer(q.nbytes)
    d_q.upload(q)
    rs = C.sizeof(sg._lib.VariantResult)
    d_r = [hip.DeviceBuffer(V * rs), hip.DeviceBuffer(V * rs)]
    pool = sg.ScoreTestPool(*args, y=nf.y, spa_pval_cutoff=cutoff, engines=3)
    one = sg.ScoreTestEngine(*args, y=nf.y, spa_pval_cutoff=cutoff)
    firsts = [int(f) for f in np.concatenate([[0], np.cumsum(sizes)[:-1]])]
    batches = [[(d_q.value + f * n, n, s, 1.0 / 127, d_r[k].value + f * rs)
                for f, s in zip(firsts, sizes)] for k in range(2)]
    pool.test_markers_u8_device(batches[0])
    hip.device_synchronize()
    one.copy_tuning_from(pool.engines[0])
    for b in batches[1]:
        one.test_markers_u8_device(*b)
    hip.device_synchronize()
    host = [np.empty(V * rs, np.uint8), np.empty(V * rs, np.uint8)]
    for k in range(2):
        d_r[k].download(host[k])
    res = [sg.results_to_dict((sg._lib.VariantResult * V).from_buffer(h)) for h in host]
    _say(f"pool of 3: {int(res[0]['is_spa'].sum())} of {V} SPA'd")
    assert res[0]["is_spa"].sum() >= 10
    np.testing.assert_array_equal(host[0], host[1])
    eo = oracle.ScoreEngine(*args, y=nf.y, spa_cutoff=cutoff)
    want = eo.test_batch_u8(q[:sizes[0]], 1.0 / 127.0)
    _compare({k: v[:sizes[0]] for k, v in res[0].items()}, want, integer_dosages=False,
             ac_rtol=5e-11)
    pool.close()
    one.close()
