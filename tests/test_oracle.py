"""Pins the CPU oracle (oracle/saige_oracle.cpp) against the reference's known-answer tests,
the golden fixtures under tests/golden/ and what the mathematics fixes.  CPU only."""
import math
import os

import mpmath
import numpy as np
import pytest
from scipy import special

import oracle
from tests.helpers import decode_np, random_bed

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------------------- decode / AF
def test_decode_kat():
    # plink.rs:289-303: codes and the packed byte 0b11_10_01_00
    for byte, want in [(0b00, 2.0), (0b10, 1.0), (0b11, 0.0)]:
        assert oracle.decode_marker(np.array([byte], np.uint8), 1)[0] == want
    assert np.isnan(oracle.decode_marker(np.array([0b01], np.uint8), 1)[0])
    d = oracle.decode_marker(np.array([0b11_10_01_00], np.uint8), 4)
    assert d[0] == 2.0 and np.isnan(d[1]) and d[2] == 1.0 and d[3] == 0.0


def test_decode_subset_matches_numpy():
    body, bpm = random_bed(37, 5, seed=1, p_missing=0.1)
    full = decode_np(body, 37)
    idx = np.array([36, 0, 5, 5, 17], np.uint32)
    for j in range(5):
        got = oracle.decode_marker(body[j], 37, idx)
        np.testing.assert_array_equal(got, full[j][idx])


def test_compute_af_frq_fixture():
    # extdata/input/plinkforGRM_1000samples_10kMarkers.frq (PLINK 1.9): AC = MAF * NCHROBS
    f = np.load(os.path.join(GOLD, "frq_subset.npz"))
    n = int(f["n"])
    for row, maf, nch in zip(f["bed"], f["maf"], f["nchrobs"]):
        s, af, mac, nv = oracle.compute_af(oracle.decode_marker(row, n))
        assert nv * 2 == nch
        assert s == round(maf * nch)
        assert mac == min(s, 2 * nv - s)


def test_compute_af_r_saige_ac():
    # extdata/output/genotype_100markers_marker_plink.txt AC_Allele2 over the 1,000 model samples
    d = np.load(os.path.join(GOLD, "ac100.npz"))
    for row, ac in zip(d["bed"], d["ac"]):
        s, af, mac, nv = oracle.compute_af(oracle.decode_marker(row, int(d["n_total"]), d["sample_idx"]))
        assert s == ac and nv == 1000


def test_compute_af_missing_and_empty():
    s, af, mac, nv = oracle.compute_af(np.array([np.nan, 2.0, 1.0, np.nan]))
    assert (s, nv) == (3.0, 2) and af == 0.75 and mac == 1.0
    s, af, mac, nv = oracle.compute_af(np.array([np.nan]))
    assert nv == 0 and af == 0.0


# --------------------------------------------------------------------------- RNG replicas
def test_chacha_core_vs_openssl():
    g = np.load(os.path.join(GOLD, "chacha20.npz"))
    for key, blocks in zip(g["keys"], g["blocks"]):
        for b in range(4):
            np.testing.assert_array_equal(oracle.chacha_block(key, b, 0, 20), blocks[b])


def test_chacha8_stream_is_8_round_keystream():
    key = oracle.seed_from_u64(12345)
    stream = oracle.rng_u32_stream(12345, 64)
    blocks = np.concatenate([oracle.chacha_block(key, b, 0, 8) for b in range(4)])
    np.testing.assert_array_equal(stream, blocks)


def test_seed_from_u64_pcg32():
    # rand_core 0.6.4 seed_from_u64: PCG32 (MUL 6364136223846793005, INC 11634580027462260723),
    # restated here in Python integers
    state = 12345
    want = []
    for _ in range(8):
        state = (state * 6364136223846793005 + 11634580027462260723) % (1 << 64)
        xs = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
        rot = state >> 59
        want.append(((xs >> rot) | (xs << ((32 - rot) % 32))) & 0xFFFFFFFF)
    np.testing.assert_array_equal(oracle.seed_from_u64(12345), np.array(want, np.uint32))


def test_pcg32_core_vs_published_reference_output():
    # pcg32-demo of the PCG reference library (pcg-c): pcg32_srandom_r(42, 54) and its first
    # six pcg32_random_r outputs, as printed by the demo ("Round 1").  Seeding: inc =
    # 2 * 54 + 1; state = 0, one step, state += 42, one step.
    mul, inc = 6364136223846793005, 2 * 54 + 1
    state = ((inc + 42) * mul + inc) % (1 << 64)
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    np.testing.assert_array_equal(oracle.pcg32_stream(state, inc, 6), np.array(want, np.uint32))


def test_seed_from_u64_is_pcg32_stream_with_rand_core_increment():
    inc = 11634580027462260723
    for seed in (0, 1, 12345, 2**64 - 1):
        start = (seed * 6364136223846793005 + inc) % (1 << 64)
        np.testing.assert_array_equal(oracle.seed_from_u64(seed), oracle.pcg32_stream(start, inc, 8))


def test_rademacher_is_sign_bit_stream():
    r = oracle.rademacher(7, 50, 3)
    s = oracle.rng_u32_stream(7, 150)
    np.testing.assert_array_equal(r.ravel(), np.where(s >> 31, 1.0, -1.0))


def test_gen_range_in_range_and_uniformish():
    v = oracle.gen_range(12345, 1000, 4085)
    assert v.max() < 4085 and len(np.unique(v)) > 850


def test_shuffle_is_permutation():
    v = oracle.shuffle(12345, 905)
    assert sorted(v.tolist()) == list(range(905))
    assert (v != np.arange(905)).mean() > 0.9


# --------------------------------------------------------------------------- GRM product
def _grm_case(n=24, m=40, seed=3, p_missing=0.05):
    body, bpm = random_bed(n, m, seed, p_missing)
    d = decode_np(body, n)
    af = np.array([oracle.compute_af(row)[1] for row in d])
    return d, af


def _dense_grm(d, af):
    m, n = d.shape
    z = np.zeros((m, n))
    for j in range(m):
        den = math.sqrt(2 * af[j] * (1 - af[j]))
        if den > 1e-10:
            g = np.where(np.isnan(d[j]), 2 * af[j], d[j])
            z[j] = (g - 2 * af[j]) / den
    return z.T @ z / m


@pytest.mark.parametrize("mode,tol", [(oracle.GRM_F64, 1e-13), (oracle.GRM_F32, 1e-5)])
def test_grm_matches_dense(mode, tol):
    d, af = _grm_case()
    K = _dense_grm(d, af)
    g = oracle.Grm(d, af, mode)
    v = np.random.default_rng(0).standard_normal(d.shape[1])
    np.testing.assert_allclose(g.mat_vec(v), K @ v, rtol=0, atol=tol * np.abs(K @ v).max())


def test_grm_symmetric_psd():
    d, af = _grm_case(n=16, m=30)
    g = oracle.Grm(d, af, oracle.GRM_F64)
    K = np.stack([g.mat_vec(e) for e in np.eye(16)], axis=1)
    np.testing.assert_allclose(K, K.T, atol=1e-13)
    assert np.linalg.eigvalsh(K).min() > -1e-12


def test_grm_monomorphic_contributes_zero():
    d, af = _grm_case(n=12, m=6, p_missing=0)
    d2 = np.vstack([d, np.full((1, 12), 2.0)])
    af2 = np.append(af, 1.0)
    v = np.arange(12.0)
    a = oracle.Grm(d, af, oracle.GRM_F64).mat_vec(v)
    b = oracle.Grm(d2, af2, oracle.GRM_F64).mat_vec(v)
    np.testing.assert_allclose(b, a * 6 / 7, rtol=1e-14)


# --------------------------------------------------------------------------- PCG
def test_pcg_kat_spd():
    a = np.array([[4.0, 2.0, 1.0], [2.0, 5.0, 3.0], [1.0, 3.0, 6.0]])
    b = np.array([1.0, 2.0, 3.0])
    x, it, res, conv = oracle.pcg_solve_dense(a, b, 1e-10, 1000)  # decomposition.rs:491-510
    assert conv and it <= 3
    np.testing.assert_allclose(a @ x, b, atol=1e-8)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-9)


def test_pcg_kat_identity_grm():
    # pcg.rs:202-223 (tau_e * W + tau_g * I with W = 1): the Sigma form 1/W + 0.1 I -> b / 1.1
    g = oracle.Grm(identity_n=5)
    b = np.arange(1.0, 6.0)
    x, it, res, conv = oracle.pcg_solve_sigma(g, np.ones(5), (1.0, 0.1), b)
    assert conv
    np.testing.assert_allclose(x, b / 1.1, atol=1e-5)


def test_pcg_sigma_vs_numpy():
    d, af = _grm_case(n=40, m=80, p_missing=0)
    K = _dense_grm(d, af)
    g = oracle.Grm(d, af, oracle.GRM_F64)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.05, 0.25, 40)
    tau = (1.0, 0.7)
    b = rng.standard_normal(40)
    x, it, res, conv = oracle.pcg_solve_sigma(g, w, tau, b, 1e-12, 500)
    S = np.diag(tau[0] / w) + tau[1] * K
    assert conv and it <= 40
    np.testing.assert_allclose(x, np.linalg.solve(S, b), rtol=1e-9)


def test_pcg_zero_rhs_and_nonconvergence():
    d, af = _grm_case(n=20, m=30, p_missing=0)
    g = oracle.Grm(d, af, oracle.GRM_F64)
    x, it, res, conv = oracle.pcg_solve_sigma(g, np.full(20, 0.2), (1.0, 2.0), np.zeros(20))
    assert conv and it == 0 and not x.any()
    b = np.random.default_rng(1).standard_normal(20)
    x, it, res, conv = oracle.pcg_solve_sigma(g, np.full(20, 0.2), (1.0, 2.0), b, 1e-14, 1)
    assert not conv and it == 1


# --------------------------------------------------------------------------- AI-REML
def _dense_ai_reml(K, y, X, probes, trait, max_iter, tol):
    """The AI-REML iteration of ai_reml.rs:135-384 with dense solves (numpy) in place of PCG."""
    n, p = X.shape
    # IRLS start (ai_reml.rs:431-482)
    mu = (y + 0.5) / 2 if trait == 0 else y.copy()
    eta = np.log(mu / (1 - mu)) if trait == 0 else mu.copy()
    alpha = np.zeros(p)
    for _ in range(25):
        w = np.maximum(mu * (1 - mu), 1e-30) if trait == 0 else np.ones(n)
        z = eta + (y - mu) / w
        an = np.linalg.solve(X.T @ (w[:, None] * X), X.T @ (w * z))
        eta = X @ an
        mu = np.clip(1 / (1 + np.exp(-eta)), 1e-10, 1 - 1e-10) if trait == 0 else eta
        ch = np.abs(an - alpha).sum()
        alpha = an
        if ch < 1e-8:
            break
    tau = np.array([1.0, 0.0])
    traj = []
    for _ in range(max_iter):
        w = np.maximum(mu * (1 - mu), 1e-30) if trait == 0 else np.ones(n)
        S = np.diag(tau[0] / w) + tau[1] * K
        Si = np.linalg.inv(S)
        yt = eta + (y - mu) / w
        SiX = Si @ X
        XSX = X.T @ SiX
        alpha = np.linalg.solve(XSX, X.T @ Si @ yt)
        eta = X @ alpha
        mu = np.clip(1 / (1 + np.exp(-eta)), 1e-10, 1 - 1e-10) if trait == 0 else eta
        P = Si - SiX @ np.linalg.solve(XSX, SiX.T)
        Py = Si @ yt - SiX @ alpha
        KPy = K @ Py
        tr = np.mean([(K @ u) @ (P @ u) for u in probes])
        score_g = Py @ KPy - tr
        ai11 = KPy @ P @ KPy
        if trait == 0:
            d1 = score_g / ai11
            step = 1.0
            tg = tau[1] + d1
            while tg < 0 and step > 1e-10:
                step *= 0.5
                tg = tau[1] + step * d1
            tg = max(tg, 1e-10)
            ch = abs(tg - tau[1]) / (abs(tg) + abs(tau[1]) + tol)
            tau = np.array([1.0, tg])
        else:
            trw = np.mean([(u / w) @ (P @ u) for u in probes])
            iwpy = Py / w
            score_e = Py @ iwpy - trw
            ai00 = iwpy @ P @ iwpy
            ai01 = iwpy @ P @ KPy
            det = ai00 * ai11 - ai01 ** 2
            d0 = (ai11 * score_e - ai01 * score_g) / det
            d1 = (-ai01 * score_e + ai00 * score_g) / det
            step = 1.0
            t0, t1 = tau[0] + d0, tau[1] + d1
            while (t0 < 0 or t1 < 0) and step > 1e-10:
                step *= 0.5
                t0, t1 = tau[0] + step * d0, tau[1] + step * d1
            t = np.array([max(t0, 1e-10), max(t1, 1e-10)])
            ch = max(abs(t[0] - tau[0]) / (abs(t[0]) + abs(tau[0]) + tol),
                     abs(t[1] - tau[1]) / (abs(t[1]) + abs(tau[1]) + tol))
            tau = t
        traj.append(tau.copy())
        if ch < tol:
            break
    return np.array(traj)


@pytest.mark.parametrize("trait", [oracle.TRAIT_BINARY, oracle.TRAIT_QUANTITATIVE])
def test_ai_reml_trajectory_vs_dense(trait):
    n, m = 120, 300
    rng = np.random.default_rng(11)
    body, _ = random_bed(n, m, seed=12)
    d = decode_np(body, n)
    af = np.array([oracle.compute_af(r)[1] for r in d])
    K = _dense_grm(d, af)
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    L = np.linalg.cholesky(K + 1e-9 * np.eye(n))
    liab = 1.5 * (L @ rng.standard_normal(n)) + 0.5 * X[:, 1] + rng.standard_normal(n)
    y = (liab > np.quantile(liab, 0.7)).astype(float) if trait == 0 else liab
    probes = oracle.rademacher(12345, n, 8)
    g = oracle.Grm(d, af, oracle.GRM_F64)
    r = oracle.fit_ai_reml(g, y, X.T, trait=trait, max_iter=8, tol=1e-6, pcg_tol=1e-13,
                           pcg_max_iter=1000, n_random_vectors=8, seed=12345)
    traj = _dense_ai_reml(K, y, X, probes, trait, 8, 1e-6)
    k = min(len(traj), len(r["tau_trace"]))
    np.testing.assert_allclose(r["tau_trace"][:k], traj[:k], rtol=1e-7, atol=1e-10)


def _reml_neg_loglik(tau, K, y, X):
    """Minus the restricted log-likelihood of y ~ N(X b, tau_e I + tau_g K), up to a constant:
    (log|S| + log|X'S^-1 X| + y'P y) / 2 with P = S^-1 - S^-1 X (X'S^-1 X)^-1 X'S^-1 (Patterson
    and Thompson 1971), evaluated through Cholesky factors."""
    te, tg = tau
    if te <= 0 or tg < 0:
        return np.inf
    n = len(y)
    S = te * np.eye(n) + tg * K
    L = np.linalg.cholesky(S)
    Si_X = np.linalg.solve(S, X)
    Si_y = np.linalg.solve(S, y)
    XSX = X.T @ Si_X
    Lx = np.linalg.cholesky(XSX)
    b = np.linalg.solve(XSX, X.T @ Si_y)
    ypy = y @ Si_y - (X.T @ Si_y) @ b
    return np.sum(np.log(np.diag(L))) + np.sum(np.log(np.diag(Lx))) + 0.5 * ypy


def test_ai_reml_converges_to_reml_optimum_exact_trace():
    """Pins the quantitative AI-REML iteration (ai_reml.rs:135-384: the score terms y'PKPy and
    y'PPy, the Hutchinson traces, the 2 x 2 average-information step, the projections) by the
    mathematics it solves: with probes sqrt(n) e_1 .. sqrt(n) e_n the Hutchinson means are the
    exact traces tr(PK) and tr(P), so the converged tau must be the maximiser of the restricted
    likelihood, which scipy finds here from the likelihood values alone (no score formula)."""
    from scipy.optimize import minimize
    n, m = 80, 200
    d, af = _grm_case(n=n, m=m, seed=41, p_missing=0.0)
    K = _dense_grm(d, af)
    rng = np.random.default_rng(5)
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    g = np.linalg.cholesky(K + 1e-9 * np.eye(n)) @ rng.standard_normal(n)
    y = X @ np.array([0.5, -0.3]) + 0.8 * g + 0.6 * rng.standard_normal(n)
    grm = oracle.Grm(d, af, oracle.GRM_F64)
    r = oracle.fit_ai_reml(grm, y, X.T, trait=oracle.TRAIT_QUANTITATIVE, max_iter=60,
                           tol=1e-10, pcg_tol=1e-12, pcg_max_iter=5000,
                           probes=np.sqrt(n) * np.eye(n))
    assert r["converged"]
    opt = minimize(_reml_neg_loglik, x0=np.array([0.5, 0.5]), args=(K, y, X),
                   method="Nelder-Mead",
                   options=dict(xatol=1e-12, fatol=1e-15, maxiter=20000, maxfev=40000))
    assert opt.success
    np.testing.assert_allclose(r["tau"], opt.x, rtol=1e-6)
    # the optimum is interior and not the start: both components moved
    assert opt.x[1] > 0.05 and abs(opt.x[0] - 1.0) > 0.05
    # a plausible mistake moves the fixed point: with the Rademacher probes (an estimated
    # trace) the converged tau is off the optimum by far more than the 1e-6 bar
    r30 = oracle.fit_ai_reml(grm, y, X.T, trait=oracle.TRAIT_QUANTITATIVE, max_iter=60,
                             tol=1e-10, pcg_tol=1e-12, pcg_max_iter=5000)
    assert np.abs(r30["tau"] - opt.x).max() / opt.x.max() > 1e-4


def test_ai_reml_identity_grm_kat():
    # ai_reml.rs:541-591: quantitative, identity GRM, tau > 0, finite negative loglik
    n = 20
    y = 1.0 + 0.1 * np.arange(n)
    g = oracle.Grm(identity_n=n)
    r = oracle.fit_ai_reml(g, y, np.ones((1, n)), trait=oracle.TRAIT_QUANTITATIVE, max_iter=10,
                           tol=1e-4, pcg_tol=1e-4, pcg_max_iter=100, n_random_vectors=5, seed=42)
    assert r["tau"][0] > 0 and r["tau"][1] >= 0
    assert 0 < r["iterations"] <= 10
    assert np.isfinite(r["log_likelihood"]) and r["log_likelihood"] < 0


# --------------------------------------------------------------------------- XVX / VR
def test_xvx_inv_xv_reference_layout():
    rng = np.random.default_rng(2)
    n, p = 9, 3
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))])
    w = rng.uniform(0.1, 0.25, n)
    A = np.linalg.inv(X.T @ (w[:, None] * X)) @ (X.T * w)  # p x n
    flat = oracle.xvx_inv_xv(X.T, w)
    np.testing.assert_allclose(flat, A.ravel(), rtol=1e-12)  # element (j,i) at j*n+i
    flat1 = oracle.xvx_inv_xv(X[:, :1].T, w)
    A1 = np.linalg.inv(X[:, :1].T @ (w[:, None] * X[:, :1])) @ (X[:, :1].T * w)
    np.testing.assert_allclose(flat1, A1.ravel(), rtol=1e-12)


def test_cal_cv_kat():
    assert oracle.cal_cv([1.0, 1.0, 1.0, 1.0]) < 1e-10   # variance_ratio.rs:351-360
    cv = oracle.cal_cv([0.9, 1.0, 1.1])
    assert 0 < cv < 1
    assert abs(cv - (0.1 / 1.0) / 3) < 1e-12
    assert oracle.cal_cv([1.0]) == math.inf


def test_select_vr_kat():
    L = oracle.lib()
    up = np.array([10.0, 20.0, np.inf])
    vr = np.array([0.8, 0.9, 1.0])
    f = lambda mac, k: L.orc_select_vr(mac, 0.95, up.ctypes.data_as(oracle._dp),
                                       vr.ctypes.data_as(oracle._dp), k)
    assert f(5.0, 3) == 0.8 and f(15.0, 3) == 0.9 and f(100.0, 3) == 1.0 and f(5.0, 0) == 0.95


def test_vr_closed_form():
    n, m = 80, 160
    body, _ = random_bed(n, m, seed=21, p_missing=0)
    d = decode_np(body, n)
    af = np.array([oracle.compute_af(r)[1] for r in d])
    K = _dense_grm(d[:120], af[:120])
    g = oracle.Grm(d[:120], af[:120], oracle.GRM_F64)
    rng = np.random.default_rng(3)
    X = np.column_stack([np.ones(n), rng.standard_normal(n)])
    mu = rng.uniform(0.05, 0.4, n)
    w = mu * (1 - mu)
    tau = (1.0, 0.3)
    a = oracle.xvx_inv_xv(X.T, w)
    vrd = d[120:]
    macs = np.array([oracle.compute_af(r)[2] for r in vrd])
    vr, used, per = oracle.estimate_vr(g, vrd, macs, mu, w, tau, X.T, a, n_markers=5, min_mac=1.0)
    # recompute per-marker ratios for the shuffled order with dense algebra
    order = oracle.shuffle(12345, len(vrd))
    Si = np.linalg.inv(np.diag(tau[0] / w) + tau[1] * K)
    Acm = a.reshape(X.shape[1], n)  # the reference's flat buffer, read col-major p x n below
    Mcm = a.reshape(n, X.shape[1]).T  # element (r, c) = flat[c*p + r]
    want = []
    for k in order:
        if macs[k] < 1.0:
            continue
        g0 = vrd[k] if vrd[k].sum() / (2 * n) <= 0.5 else 2 - vrd[k]
        ac = g0.sum()
        gt = g0 - X @ (Mcm @ g0)
        want.append((gt @ Si @ gt / ac) / np.sum(mu * (1 - mu) * gt * gt / ac))
        if len(want) == used:
            break
    del Acm
    np.testing.assert_allclose(per, want, rtol=1e-6)
    assert abs(vr - np.mean(want)) < 1e-6 * vr


# --------------------------------------------------------------------------- score test
def test_case_control_af_kat():
    # single_variant.rs:386-402
    n = 10
    y = np.array([1.0] * 5 + [0.0] * 5)
    e = oracle.ScoreEngine(np.full(n, 0.5), y - 0.5, (1.0, 0.1), np.ones((1, n)),
                           np.full(n, 1.0 / n), 1.0, y=y, use_spa=False)
    r = e.test_marker(np.array([0.0, 1, 2, 0, 1, 0, 2, 1, 0, 0]))
    assert r["n_cases"] == 5 and r["n_controls"] == 5
    assert abs(r["af_cases"] - 0.4) < 1e-10 and abs(r["af_controls"] - 0.3) < 1e-10


def _engine_case(n=200, p=2, seed=4):
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(p - 1)])
    mu = 1 / (1 + np.exp(-(-2.0 + 0.3 * X[:, -1])))
    y = (rng.random(n) < mu).astype(float)
    a = oracle.xvx_inv_xv(X.T, mu * (1 - mu))
    return X, mu, y, a


def test_score_closed_form():
    X, mu, y, a = _engine_case()
    n, p = X.shape
    e = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.1, y=y, use_spa=False)
    g = np.random.default_rng(1).binomial(2, 0.2, n).astype(float)
    r = e.test_marker(g)
    M = a.reshape(n, p).T
    gt = g - X @ (M @ g)
    S = gt @ (y - mu)
    vt = np.sum(gt * gt * mu * (1 - mu))
    assert abs(r["score"] - S) < 1e-10 * abs(S) + 1e-12
    assert abs(r["var_t"] - vt) < 1e-12 * vt
    T = S * S / (vt * 1.1)
    assert abs(r["pvalue"] - special.chdtrc(1, T)) < 1e-12
    assert abs(r["beta"] - S / (vt * 1.1)) < 1e-10 * abs(r["beta"])


def test_score_allele_flip_and_monomorphic():
    # property_tests.rs:359-413 (p and |beta| invariant under g -> 2 - g for p = 1) and
    # :312-349 (monomorphic -> no NaN)
    X, mu, y, a = _engine_case(p=1)
    e = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y, use_spa=True)
    g = np.random.default_rng(2).binomial(2, 0.1, len(mu)).astype(float)
    r1, r2 = e.test_marker(g), e.test_marker(2 - g)
    assert abs(r1["pvalue"] - r2["pvalue"]) < 1e-10
    assert abs(abs(r1["beta"]) - abs(r2["beta"])) < 1e-10
    r0 = e.test_marker(np.zeros(len(mu)))
    assert not np.isnan(r0["pvalue"]) and r0["pvalue"] == 1.0


def test_score_missing_dosage_reference_semantics():
    X, mu, y, a = _engine_case(p=1)
    e = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X.T, a, 1.0, y=y)
    g = np.random.default_rng(2).binomial(2, 0.2, len(mu)).astype(float)
    g[3] = np.nan
    r = e.test_marker(g)
    assert np.isnan(r["score"]) and r["tstat"] == 0.0 and r["pvalue"] == 1.0
    assert r["beta"] == 0.0 and math.isinf(r["se_beta"]) and r["n"] == len(mu) - 1


# --------------------------------------------------------------------------- SPA
def test_cgf_kats():
    mu = np.array([0.3, 0.5, 0.7])
    g = np.array([0.0, 1.0, 2.0])
    assert abs(oracle.k0(0.0, mu, g)) < 1e-10          # binary.rs:246-252
    assert abs(oracle.k1_adj(0.0, mu, g, 0.0) - 1.9) < 1e-10
    assert oracle.k2(0.0, mu, g) > 0


def test_cgf_derivatives_finite_difference():
    rng = np.random.default_rng(9)
    mu = rng.uniform(0.01, 0.3, 30)
    g = rng.standard_normal(30)
    mpmath.mp.dps = 40
    K = lambda t: mpmath.fsum(mpmath.log(1 - m + m * mpmath.exp(gi * t)) for m, gi in zip(mu, g))
    for t in (-0.7, 0.0, 0.4, 1.3):
        d1 = float(mpmath.diff(K, t))
        d2 = float(mpmath.diff(K, t, 2))
        assert abs(oracle.k1_adj(t, mu, g, 0.0) - d1) < 1e-12 * max(1, abs(d1))
        assert abs(oracle.k2(t, mu, g) - d2) < 1e-12 * max(1, abs(d2))
        assert abs(oracle.k0(t, mu, g) - float(K(t))) < 1e-12 * max(1, abs(float(K(t))))


def test_root_converges_and_extreme_q():
    mu = np.array([0.3, 0.5, 0.7, 0.4, 0.6])
    g = np.array([0.0, 1.0, 2.0, 0.0, 1.0])
    r = oracle.spa_binary(mu, g, 1.0, 2 * (mu @ g) - 1.0, 0.04, fast=False)
    assert r["converged"]                                 # binary.rs:272-279
    # the root solves K'(t) = q
    assert abs(oracle.k1_adj(r["root1"], mu, g, 1.0)) < 1e-6
    r = oracle.spa_binary(np.full(10, 0.5), np.ones(10), 20.0, -15.0, 0.01)
    assert r["converged"] and math.isinf(r["root1"])       # binary.rs:300-307


def test_spa_pvalue_vs_exact_tail():
    # Lugannani-Rice vs the exact distribution of sum g_i Y_i (enumeration, n = 14):
    # accuracy of the method, not parity.
    rng = np.random.default_rng(5)
    n = 14
    mu = rng.uniform(0.05, 0.3, n)
    g = rng.integers(0, 3, n).astype(float)
    probs = {0.0: 1.0}
    for m, gi in zip(mu, g):
        nxt = {}
        for s, pr in probs.items():
            nxt[s] = nxt.get(s, 0) + pr * (1 - m)
            nxt[s + gi] = nxt.get(s + gi, 0) + pr * m
        probs = nxt
    m1 = mu @ g
    q = m1 + 2.5 * math.sqrt(np.sum(mu * (1 - mu) * g * g))
    exact = sum(pr for s, pr in probs.items() if s >= q - 1e-12) + \
        sum(pr for s, pr in probs.items() if s <= 2 * m1 - q + 1e-12)
    r = oracle.spa_binary(mu, g, q, 2 * m1 - q, 0.01, fast=False)
    assert r["converged"] and 0 <= r["pvalue"] <= 1
    assert abs(math.log(r["pvalue"]) - math.log(exact)) < 0.5


def test_spa_fast_equals_full_on_adjusted_genotypes():
    # with covariate-adjusted g~ no entry is exactly 0, so fast == full (fast.rs:185-204)
    rng = np.random.default_rng(6)
    mu = rng.uniform(0.02, 0.2, 300)
    g = rng.binomial(2, 0.05, 300) - 0.1
    m1 = mu @ g
    q = m1 + 3 * math.sqrt(np.sum(mu * (1 - mu) * g * g))
    a = oracle.spa_binary(mu, g, q, 2 * m1 - q, 0.003, fast=True)
    b = oracle.spa_binary(mu, g, q, 2 * m1 - q, 0.003, fast=False)
    assert a["pvalue"] == b["pvalue"]


def test_pvalue_closed_forms():
    for t in (0.0, 0.5, 4.0, 20.0, 45.0):
        assert abs(oracle.chisq1_sf(t) - special.chdtrc(1, t)) < 2e-16 + 1e-14 * special.chdtrc(1, t) * 1e2
    for z in (-5.0, -1.0, 0.0, 2.0, 6.0):
        assert abs(oracle.norm_cdf(z) - special.ndtr(z)) < 1e-16 + 1e-15 * special.ndtr(z)


def test_categorical_vr_bins():
    """compute_categorical_vr (variance_ratio.rs:263-287): a marker is in the bin whose lower
    bound it exceeds and whose upper bound it does not; empty bins are left out; the ratio of
    a bin is the mean over its markers, in the order they were used."""
    vr = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    mac = [1.0, 1.5, 2.5, 20.5, 20.6, 1000.0]
    got = oracle.categorical_vr(vr, mac)
    assert got == [(1.5, 1.5), (2.5, 3.0), (20.5, 4.0), (float("inf"), 5.5)]
    assert oracle.categorical_vr([], []) == []
    assert oracle.categorical_vr([0.9], [0.5]) == []  # MAC 0.5 is in no bin (lower bound excluded)
