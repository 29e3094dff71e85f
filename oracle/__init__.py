"""CPU oracle — test infrastructure only.

ctypes bindings for ``liboracle_saige.so``, the C++ restatement of the saige-rs CPU path
(``oracle/saige_oracle.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the checker or
the reported CPU baseline.  The product (``csaige_amd``) never imports it.

Functions pinned only by the mathematics or by invariants are marked in their tests;
the ChaCha core is pinned by OpenSSL's chacha20 keystream and PCG32 (seed_from_u64) by the
PCG reference library's published output; the rand 0.8.5 draws (gen::<bool>, gen_range,
shuffle) and the statrs CDF restatements are "parity unpinned" beyond those checks (no
reference test pins their draws or last bits).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle_saige.so")

GRM_F32 = 0  # saige-rs OnTheFlyGrm (f32 storage + f32 accumulation), pcg.rs:88-180
GRM_F64 = 1  # same algorithm in f64: the FP64 oracle

TRAIT_BINARY = 0
TRAIT_QUANTITATIVE = 1


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def _load():
    if not os.path.exists(_SO):
        build()
    return C.CDLL(_SO)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)


class RemlConfig(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("pcg_tol", C.c_double),
                ("pcg_max_iter", C.c_int), ("n_random_vectors", C.c_int), ("seed", C.c_uint64)]


class RemlResult(C.Structure):
    _fields_ = [("tau", C.c_double * 2), ("iterations", C.c_int), ("converged", C.c_int),
                ("log_likelihood", C.c_double), ("grm_products", C.c_longlong),
                ("pcg_iterations", C.c_longlong)]


class FitNullOut(C.Structure):
    _fields_ = [("tau", C.c_double * 2), ("iterations", C.c_int), ("converged", C.c_int),
                ("vr", C.c_double), ("vr_markers_used", C.c_int), ("n_grm_markers", C.c_int),
                ("n_vr_markers", C.c_int), ("grm_products", C.c_longlong),
                ("pcg_iterations", C.c_longlong), ("log_likelihood", C.c_double)]


class SpaResult(C.Structure):
    _fields_ = [("pvalue", C.c_double), ("converged", C.c_int), ("is_spa", C.c_int),
                ("iters1", C.c_int), ("iters2", C.c_int), ("root1", C.c_double),
                ("root2", C.c_double)]


class Engine(C.Structure):
    _fields_ = [("mu", _dp), ("mu2", _dp), ("res", _dp), ("tau_e", C.c_double),
                ("tau_g", C.c_double), ("a", _dp), ("x", _dp), ("n", C.c_size_t),
                ("p", C.c_size_t), ("vr", C.c_double), ("cat_upper", _dp), ("cat_vr", _dp),
                ("n_cat", C.c_size_t), ("use_spa", C.c_int), ("use_fast_spa", C.c_int),
                ("spa_tol", C.c_double), ("spa_cutoff", C.c_double), ("y", _dp)]


class VariantResult(C.Structure):
    _fields_ = [("ac", C.c_double), ("af", C.c_double), ("mac", C.c_double), ("n", C.c_uint64),
                ("score", C.c_double), ("var_t", C.c_double), ("var_t_star", C.c_double),
                ("tstat", C.c_double), ("pvalue", C.c_double), ("pvalue_noadj", C.c_double),
                ("is_spa", C.c_int), ("spa_converged", C.c_int), ("beta", C.c_double),
                ("se_beta", C.c_double), ("vr", C.c_double), ("af_cases", C.c_double),
                ("af_controls", C.c_double), ("n_cases", C.c_uint64),
                ("n_controls", C.c_uint64), ("spa_iters1", C.c_int), ("spa_iters2", C.c_int)]


def _declare(L):
    L.orc_decode_marker.argtypes = [_u8p, C.c_size_t, _u32p, C.c_size_t, _dp]
    L.orc_bgen_layout2_dosages.argtypes = [_u8p, C.c_size_t, _dp, C.c_size_t,
                                           C.POINTER(C.c_size_t)]
    L.orc_compute_af.argtypes = [_dp, C.c_size_t, _dp, _dp, _dp, _u64p]
    L.orc_chacha_block.argtypes = [_u32p, C.c_uint64, C.c_uint64, C.c_int, _u32p]
    L.orc_seed_from_u64.argtypes = [C.c_uint64, _u32p]
    L.orc_pcg32_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, _u32p]
    L.orc_rng_u32_stream.argtypes = [C.c_uint64, C.c_size_t, _u32p]
    L.orc_rademacher.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, _dp]
    L.orc_gen_range.argtypes = [C.c_uint64, C.c_size_t, C.c_uint64, _u64p]
    L.orc_shuffle.argtypes = [C.c_uint64, C.c_size_t, _u64p]
    L.orc_grm_new.argtypes = [_dp, C.c_size_t, C.c_size_t, _dp, C.c_int]
    L.orc_grm_new.restype = C.c_void_p
    L.orc_grm_identity.argtypes = [C.c_size_t]
    L.orc_grm_identity.restype = C.c_void_p
    L.orc_grm_free.argtypes = [C.c_void_p]
    L.orc_grm_mat_vec.argtypes = [C.c_void_p, _dp, _dp]
    L.orc_grm_mat_vec_threads.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    L.orc_pcg_solve_sigma.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp, C.c_size_t,
                                      C.c_double, C.c_int, _dp, _ip, _dp]
    L.orc_pcg_solve_dense.argtypes = [_dp, C.c_size_t, _dp, C.c_double, C.c_int, _dp, _ip, _dp]
    L.orc_fit_ai_reml.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, C.c_size_t, C.c_int,
                                  C.POINTER(RemlConfig), C.POINTER(RemlResult), _dp, _dp, _dp,
                                  _dp, _dp, _dp]
    L.orc_fit_ai_reml_probes.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, C.c_size_t, C.c_int,
                                         C.POINTER(RemlConfig), C.POINTER(RemlResult), _dp, _dp,
                                         _dp, _dp, _dp, _dp, _dp]
    L.orc_xvx_inv_xv.argtypes = [_dp, C.c_size_t, C.c_size_t, _dp, _dp]
    L.orc_cal_cv.argtypes = [_dp, C.c_size_t]
    L.orc_cal_cv.restype = C.c_double
    L.orc_estimate_vr.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, _dp, _dp, _dp, C.c_int, _dp,
                                  _dp, C.c_size_t, C.c_size_t, C.c_int, C.c_double, C.c_double,
                                  C.c_uint64, _dp, _ip, _dp, _llp]
    L.orc_fit_null.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u32p, C.c_size_t, _dp,
                               _dp, C.c_size_t, C.c_int, C.POINTER(RemlConfig), C.c_double,
                               C.c_double, C.c_int, C.c_int, C.POINTER(FitNullOut), _dp, _dp,
                               _u8p, _dp, _dp, _dp]
    for nm in ("orc_k0_binom", "orc_k2_binom"):
        getattr(L, nm).argtypes = [C.c_double, _dp, _dp, C.c_size_t]
        getattr(L, nm).restype = C.c_double
    L.orc_k1_adj_binom.argtypes = [C.c_double, _dp, _dp, C.c_size_t, C.c_double]
    L.orc_k1_adj_binom.restype = C.c_double
    L.orc_spa_binary.argtypes = [_dp, _dp, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_double, C.c_double,
                                 C.POINTER(SpaResult)]
    L.orc_select_vr.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_size_t]
    L.orc_select_vr.restype = C.c_double
    L.orc_test_marker.argtypes = [C.POINTER(Engine), _dp, C.POINTER(VariantResult)]
    L.orc_test_batch_u8.argtypes = [C.POINTER(Engine), _u8p, C.c_size_t, C.c_size_t, C.c_double,
                                    C.POINTER(VariantResult)]
    L.orc_chisq1_sf.argtypes = [C.c_double]
    L.orc_chisq1_sf.restype = C.c_double
    L.orc_norm_cdf.argtypes = [C.c_double]
    L.orc_norm_cdf.restype = C.c_double


# ----------------------------------------------------------------------------------
# numpy helpers
# ----------------------------------------------------------------------------------
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(t)


def decode_marker(marker_bytes, n_total, sample_idx=None):
    """plink.rs:170-267 — one marker's 2-bit codes to f64 dosages (NaN = missing)."""
    mb = np.ascontiguousarray(marker_bytes, dtype=np.uint8)
    if sample_idx is None:
        n = n_total
        sp = None
    else:
        si = np.ascontiguousarray(sample_idx, dtype=np.uint32)
        n = len(si)
        sp = _p(si, _u32p)
    out = np.empty(n, np.float64)
    lib().orc_decode_marker(_p(mb, _u8p), n_total, sp, n, _p(out, _dp))
    return out


def decode_bed(bed_body, bytes_per_marker, n_total, m, sample_idx=None):
    """Decode m markers; returns n x m (col-major as an (m, n) C array's transpose)."""
    body = np.frombuffer(bytes(bed_body), dtype=np.uint8) if not isinstance(bed_body, np.ndarray) else bed_body
    cols = [decode_marker(body[j * bytes_per_marker:(j + 1) * bytes_per_marker], n_total, sample_idx)
            for j in range(m)]
    return np.stack(cols, axis=0)  # (m, n): row j = marker j


def compute_af(d):
    """traits.rs:42-54 — (sum, af, mac, n_valid)."""
    d = _f64(d)
    s, af, mac = C.c_double(), C.c_double(), C.c_double()
    nv = C.c_uint64()
    lib().orc_compute_af(_p(d, _dp), len(d), C.byref(s), C.byref(af), C.byref(mac), C.byref(nv))
    return s.value, af.value, mac.value, nv.value


def chacha_block(key, counter, stream=0, rounds=8):
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(16, np.uint32)
    lib().orc_chacha_block(_p(k, _u32p), counter, stream, rounds, _p(out, _u32p))
    return out


def seed_from_u64(seed):
    k = np.empty(8, np.uint32)
    lib().orc_seed_from_u64(seed, _p(k, _u32p))
    return k


def pcg32_stream(state, inc, count):
    """PCG32 outputs (XSH-RR of the state before each step) from `state` with increment
    `inc`; seed_from_u64(x) is this stream from x * 6364136223846793005 + INC."""
    out = np.empty(count, np.uint32)
    lib().orc_pcg32_stream(state, inc, count, _p(out, _u32p))
    return out


def rng_u32_stream(seed, count):
    out = np.empty(count, np.uint32)
    lib().orc_rng_u32_stream(seed, count, _p(out, _u32p))
    return out


def rademacher(seed, n, nprobe):
    """ai_reml.rs:121-130 — returns (nprobe, n)."""
    out = np.empty((nprobe, n), np.float64)
    lib().orc_rademacher(seed, n, nprobe, _p(out, _dp))
    return out


def gen_range(seed, count, n):
    out = np.empty(count, np.uint64)
    lib().orc_gen_range(seed, count, n, _p(out, _u64p))
    return out


def shuffle(seed, length):
    out = np.empty(length, np.uint64)
    lib().orc_shuffle(seed, length, _p(out, _u64p))
    return out


class Grm:
    """OnTheFlyGrm (pcg.rs:91-181).  ``dosages`` is (m, n): one row per marker."""

    def __init__(self, dosages=None, afs=None, mode=GRM_F64, identity_n=None):
        L = lib()
        if identity_n is not None:
            self.n = identity_n
            self.h = L.orc_grm_identity(identity_n)
            return
        d = _f64(dosages)
        m, n = d.shape
        self.n = n
        self._keep = (d, _f64(afs))
        self.h = L.orc_grm_new(_p(d, _dp), n, m, _p(self._keep[1], _dp), mode)

    def mat_vec(self, v):
        v = _f64(v)
        out = np.empty(self.n, np.float64)
        lib().orc_grm_mat_vec(self.h, _p(v, _dp), _p(out, _dp))
        return out

    def mat_vec_threads(self, v, threads):
        """The f32 product split as saige-rs splits it over a rayon pool (timing)."""
        v = _f64(v)
        out = np.empty(self.n, np.float64)
        lib().orc_grm_mat_vec_threads(self.h, _p(v, _dp), _p(out, _dp), int(threads))
        return out

    def __del__(self):
        try:
            lib().orc_grm_free(self.h)
        except Exception:
            pass


def pcg_solve_sigma(grm, w, tau, b, tol=1e-5, max_iter=500):
    w, b = _f64(w), _f64(b)
    x = np.empty(len(b), np.float64)
    it, res = C.c_int(), C.c_double()
    conv = lib().orc_pcg_solve_sigma(grm.h, _p(w, _dp), tau[0], tau[1], _p(b, _dp), len(b), tol,
                                     max_iter, _p(x, _dp), C.byref(it), C.byref(res))
    return x, it.value, res.value, bool(conv)


def pcg_solve_dense(a, b, tol=1e-5, max_iter=500):
    a, b = _f64(a), _f64(b)
    x = np.empty(len(b), np.float64)
    it, res = C.c_int(), C.c_double()
    conv = lib().orc_pcg_solve_dense(_p(a, _dp), len(b), _p(b, _dp), tol, max_iter, _p(x, _dp),
                                     C.byref(it), C.byref(res))
    return x, it.value, res.value, bool(conv)


def fit_ai_reml(grm, y, x_cm, trait=TRAIT_BINARY, max_iter=30, tol=1e-5, pcg_tol=1e-5,
                pcg_max_iter=500, n_random_vectors=30, seed=12345, probes=None):
    """ai_reml.rs:86-427.  ``x_cm`` is (p, n): row j = covariate column j.  ``probes``
    ((R, n), test use only) replaces the ChaCha8 Rademacher probes, e.g. sqrt(n) times the
    unit vectors for the exact trace."""
    y = _f64(y)
    x = _f64(x_cm)
    p, n = x.shape
    if probes is not None:
        probes = _f64(probes)
        n_random_vectors = probes.shape[0]
    cfg = RemlConfig(max_iter, tol, pcg_tol, pcg_max_iter, n_random_vectors, seed)
    res = RemlResult()
    mu, w, r, eta = (np.empty(n) for _ in range(4))
    alpha = np.empty(p)
    trace = np.full(2 * max_iter, np.nan)
    args = (grm.h, _p(y, _dp), _p(x, _dp), n, p, trait, C.byref(cfg), C.byref(res), _p(mu, _dp),
            _p(w, _dp), _p(r, _dp), _p(eta, _dp), _p(alpha, _dp), _p(trace, _dp))
    if probes is None:
        rc = lib().orc_fit_ai_reml(*args)
    else:
        rc = lib().orc_fit_ai_reml_probes(*args, _p(probes, _dp))
    if rc != 0:
        raise RuntimeError(f"orc_fit_ai_reml failed: {rc}")
    return dict(tau=np.array(res.tau[:]), iterations=res.iterations, converged=bool(res.converged),
                log_likelihood=res.log_likelihood, grm_products=res.grm_products,
                pcg_iterations=res.pcg_iterations, mu=mu, w=w, residuals=r, eta=eta, alpha=alpha,
                tau_trace=trace[:2 * res.iterations].reshape(-1, 2))


def xvx_inv_xv(x_cm, w):
    """fit_null.rs:289-307 — returns the flat p*n buffer exactly as the reference fills it."""
    x = _f64(x_cm)
    p, n = x.shape
    w = _f64(w)
    out = np.empty(p * n)
    if lib().orc_xvx_inv_xv(_p(x, _dp), n, p, _p(w, _dp), _p(out, _dp)) != 0:
        raise RuntimeError("XVX not positive definite")
    return out


def categorical_vr(per_marker_vr, mac):
    """compute_categorical_vr (saige-core/src/glmm/variance_ratio.rs:263-287) with the default
    MAC bins of VarianceRatioConfig (:50-51): (upper bound, mean ratio) of each bin that holds
    a marker, in bin order; a marker is in a bin when lower < MAC <= upper."""
    lo = [0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 10.5, 20.5]
    hi = [1.5, 2.5, 3.5, 4.5, 5.5, 10.5, 20.5, float("inf")]
    out = []
    for a, b in zip(lo, hi):
        vals = [v for v, m in zip(per_marker_vr, mac) if a < m <= b]
        if vals:
            s = 0.0
            for v in vals:
                s += v
            out.append((b, s / len(vals)))
    return out


def cal_cv(v):
    v = _f64(v)
    return lib().orc_cal_cv(_p(v, _dp), len(v))


def estimate_vr(grm, dosages, macs, mu, w, tau, x_cm, a_flat, trait=TRAIT_BINARY, n_markers=30,
                min_mac=20.0, cv_cutoff=1e-3, seed=12345):
    d = _f64(dosages)
    m, n = d.shape
    x = _f64(x_cm)
    p = x.shape[0]
    macs, mu, w, a = _f64(macs), _f64(mu), _f64(w), _f64(a_flat)
    t = _f64(tau)
    vr, used = C.c_double(), C.c_int()
    per = np.empty(m)
    prods = C.c_longlong()
    rc = lib().orc_estimate_vr(grm.h, _p(d, _dp), _p(macs, _dp), m, _p(mu, _dp), _p(w, _dp),
                               _p(t, _dp), trait, _p(x, _dp), _p(a, _dp), n, p, n_markers,
                               min_mac, cv_cutoff, seed, C.byref(vr), C.byref(used),
                               _p(per, _dp), C.byref(prods))
    if rc != 0:
        raise RuntimeError(f"orc_estimate_vr failed: {rc}")
    return vr.value, used.value, per[:used.value]


def run_loco(y, x_cm, marker_chroms, dosages, afs, trait=TRAIT_BINARY, grm_mode=GRM_F64,
             **reml):
    """run_loco (saige-core/src/glmm/loco.rs:54-114): for every chromosome, in sorted order
    of the names, fit_ai_reml on the OnTheFlyGrm of the markers on the other chromosomes.
    Returns {chromosome: fit_ai_reml result}."""
    chroms = sorted(set(marker_chroms))
    d = _f64(dosages)
    a = _f64(afs)
    out = {}
    for c in chroms:
        keep = np.array([mc != c for mc in marker_chroms])
        g = Grm(d[keep], a[keep], grm_mode)
        out[c] = fit_ai_reml(g, y, x_cm, trait=trait, **reml)
    return out


def bgen_layout2_dosages(block):
    """parse_layout2_probabilities (bgen.rs:373-430) of one decompressed block."""
    b = np.frombuffer(bytes(block), dtype=np.uint8)
    cap = int(np.frombuffer(b[:4].tobytes(), "<u4")[0]) if len(b) >= 4 else 0
    out = np.empty(max(cap, 1))
    n = C.c_size_t(0)
    rc = lib().orc_bgen_layout2_dosages(_p(np.ascontiguousarray(b), _u8p), len(b), _p(out, _dp),
                                        cap, C.byref(n))
    if rc == -1:
        raise ValueError("BGEN: 0 bits per probability not supported")
    if rc != 0:
        raise ValueError("BGEN: malformed layout-2 block")
    return out[:n.value].copy()


def fit_null(bed_body, bytes_per_marker, n_total, m_total, sample_idx, y, x_cm,
             trait=TRAIT_BINARY, grm_mode=GRM_F64, max_iter=30, tol=1e-5, seed=12345,
             n_random_vectors=30, min_maf=0.01, max_missing=0.15, n_markers_vr=30):
    """fit_null.rs:199-394 (QC, VR split, OnTheFlyGrm, AI-REML, XVX^-1X'V, VR)."""
    body = np.ascontiguousarray(bed_body, dtype=np.uint8)
    si = np.ascontiguousarray(sample_idx, dtype=np.uint32)
    n = len(si)
    y = _f64(y)
    x = _f64(x_cm)
    p = x.shape[0]
    cfg = RemlConfig(max_iter, tol, 1e-5, 500, n_random_vectors, seed)
    out = FitNullOut()
    mu = np.empty(n)
    a = np.empty(p * n)
    cls = np.empty(m_total, np.uint8)
    trace = np.full(2 * max_iter, np.nan)
    alpha = np.empty(p)
    per = np.empty(max(m_total, 1))
    rc = lib().orc_fit_null(_p(body, _u8p), bytes_per_marker, n_total, m_total, _p(si, _u32p), n,
                            _p(y, _dp), _p(x, _dp), p, trait, C.byref(cfg), min_maf, max_missing,
                            n_markers_vr, grm_mode, C.byref(out), _p(mu, _dp), _p(a, _dp),
                            _p(cls, _u8p), _p(trace, _dp), _p(alpha, _dp), _p(per, _dp))
    if rc != 0:
        raise RuntimeError(f"orc_fit_null failed: {rc}")
    return dict(tau=np.array(out.tau[:]), iterations=out.iterations, converged=bool(out.converged),
                vr=out.vr, vr_markers_used=out.vr_markers_used, n_grm_markers=out.n_grm_markers,
                n_vr_markers=out.n_vr_markers, grm_products=out.grm_products,
                pcg_iterations=out.pcg_iterations, log_likelihood=out.log_likelihood, mu=mu,
                xvx_inv_xv=a, marker_class=cls, alpha=alpha,
                per_marker_vr=per[:out.vr_markers_used].copy(),
                tau_trace=trace[:2 * out.iterations].reshape(-1, 2))


def k0(t, mu, g):
    mu, g = _f64(mu), _f64(g)
    return lib().orc_k0_binom(t, _p(mu, _dp), _p(g, _dp), len(mu))


def k1_adj(t, mu, g, q):
    mu, g = _f64(mu), _f64(g)
    return lib().orc_k1_adj_binom(t, _p(mu, _dp), _p(g, _dp), len(mu), q)


def k2(t, mu, g):
    mu, g = _f64(mu), _f64(g)
    return lib().orc_k2_binom(t, _p(mu, _dp), _p(g, _dp), len(mu))


def spa_binary(mu, g, q, qinv, pval_noadj, tol=1e-6, fast=True, init1=0.0, init2=0.0):
    mu, g = _f64(mu), _f64(g)
    r = SpaResult()
    lib().orc_spa_binary(_p(mu, _dp), _p(g, _dp), len(mu), q, qinv, pval_noadj, tol, int(fast),
                         init1, init2, C.byref(r))
    return dict(pvalue=r.pvalue, converged=bool(r.converged), is_spa=bool(r.is_spa),
                iters1=r.iters1, iters2=r.iters2, root1=r.root1, root2=r.root2)


class ScoreEngine:
    """ScoreTestEngine (single_variant.rs:77-108) over numpy arrays."""

    def __init__(self, mu, residuals, tau, x_cm, a_flat, vr, y=None, use_spa=True,
                 use_fast_spa=True, spa_tol=1e-6, spa_cutoff=0.05, cat_vr=()):
        self.mu = _f64(mu)
        self.mu2 = self.mu * (1.0 - self.mu)
        self.res = _f64(residuals)
        self.x = _f64(x_cm)
        self.a = _f64(a_flat)
        self.y = None if y is None else _f64(y)
        self.cat_upper = _f64([c[0] for c in cat_vr]) if cat_vr else np.zeros(1)
        self.cat_vr = _f64([c[1] for c in cat_vr]) if cat_vr else np.zeros(1)
        p, n = self.x.shape
        self.e = Engine(_p(self.mu, _dp), _p(self.mu2, _dp), _p(self.res, _dp), tau[0], tau[1],
                        _p(self.a, _dp), _p(self.x, _dp), n, p, vr, _p(self.cat_upper, _dp),
                        _p(self.cat_vr, _dp), len(cat_vr), int(use_spa), int(use_fast_spa),
                        spa_tol, spa_cutoff, None if self.y is None else _p(self.y, _dp))
        self.n = n

    def test_marker(self, g):
        g = _f64(g)
        r = VariantResult()
        lib().orc_test_marker(C.byref(self.e), _p(g, _dp), C.byref(r))
        return {f: getattr(r, f) for f, _ in VariantResult._fields_}

    def test_batch_u8(self, q, scale):
        q = np.ascontiguousarray(q, dtype=np.uint8)
        nvar, ld = q.shape
        out = (VariantResult * nvar)()
        lib().orc_test_batch_u8(C.byref(self.e), _p(q, _u8p), ld, nvar, scale, out)
        return results_to_numpy(out)


def results_to_numpy(arr):
    names = [f for f, _ in VariantResult._fields_]
    return {f: np.array([getattr(r, f) for r in arr]) for f in names}


def chisq1_sf(t):
    return lib().orc_chisq1_sf(t)


def norm_cdf(z):
    return lib().orc_norm_cdf(z)
