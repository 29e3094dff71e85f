"""csaige_amd — MI355X (gfx950) implementation of saige-rs's fit-null and step-2 hot paths.

This module is the Python host mirror of the saige-rs API for those paths, over the C ABI
of ``include/saige_gpu.h`` (libsaige_gpu.so):

    OnTheFlyGrm          saige-core/src/glmm/pcg.rs:91-181        (packed genotypes in HBM)
    solve_sigma          saige-linalg/src/decomposition.rs:212-298 with ai_reml.rs:142-164
    fit_ai_reml          saige-core/src/glmm/ai_reml.rs:86-427
    xvx_inv_xv           saige-cli/src/commands/fit_null.rs:289-307
    estimate_variance_ratio  saige-core/src/glmm/variance_ratio.rs:87-246
    fit_null             saige-cli/src/commands/fit_null.rs:199-394
    ScoreTestEngine      saige-core/src/score_test/single_variant.rs:77-257
    spa_binary_fast      saige-core/src/spa/fast.rs:207-242

Every call runs on the GPU; a missing library raises ImportError (no CPU fallback).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (PREC_FP64, PREC_MIXED, PREC_INT8, PREC_INT8X2, PREC_F32, TRAIT_BINARY, TRAIT_QUANTITATIVE, TRAIT_SURVIVAL,
                   SaigeError, check, load)

__all__ = ["OnTheFlyGrm", "fit_ai_reml", "solve_sigma", "xvx_inv_xv", "estimate_variance_ratio",
           "fit_null", "fit_null_multi", "ScoreTestEngine", "ScoreTestPool", "spa_binary_fast", "marker_stats", "decode",
           "PREC_FP64", "PREC_MIXED", "PREC_INT8", "PREC_INT8X2", "PREC_F32", "TRAIT_BINARY", "TRAIT_QUANTITATIVE", "TRAIT_SURVIVAL",
           "SaigeError", "load", "results_to_dict", "SpaRootMlp", "spa_training_set",
           "launch_log_enable", "launch_log_summary", "pcg_graphs",
           "pcg_graph_counts"]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(t)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _idx(sample_idx):
    if sample_idx is None:
        return None, None
    si = np.ascontiguousarray(sample_idx, dtype=np.uint32)
    return si, _p(si, _lib.u32p)


def marker_stats(bed_body, bytes_per_marker, n_total, m, sample_idx=None, device=0):
    """read_marker + compute_af (plink.rs:237-267, traits.rs:42-54) for m markers.
    Returns (ac, n_valid, af, mac)."""
    L = load()
    body = _u8(bed_body)
    si, sp = _idx(sample_idx)
    n = n_total if si is None else len(si)
    ac, af, mac = np.empty(m), np.empty(m), np.empty(m)
    nv = np.empty(m, np.uint32)
    check(L.saige_plink_marker_stats(device, _p(body, _lib.u8p), bytes_per_marker, n_total, m, sp,
                                     n, _p(ac, _lib.dp), _p(nv, _lib.u32p), _p(af, _lib.dp),
                                     _p(mac, _lib.dp)))
    return ac, nv, af, mac


def decode(bed_body, bytes_per_marker, n_total, m, sample_idx=None, device=0):
    """Dosages (m, n) with NaN for missing (plink.rs:170-267)."""
    L = load()
    body = _u8(bed_body)
    si, sp = _idx(sample_idx)
    n = n_total if si is None else len(si)
    out = np.empty((m, n))
    check(L.saige_plink_decode(device, _p(body, _lib.u8p), bytes_per_marker, n_total, m, sp, n,
                               _p(out, _lib.dp)))
    return out


class OnTheFlyGrm:
    """OnTheFlyGrm (pcg.rs:91-181) on packed genotypes resident in HBM.

    ``bed_body``: SNP-major .bed body without magic; ``marker_idx``: markers used (file
    indices, or None for all m rows); ``af``: compute_af of each used marker over the
    sample subset.
    """

    def __init__(self, bed_body, bytes_per_marker, n_total, af, sample_idx=None, marker_idx=None,
                 precision=PREC_FP64, device=0):
        L = load()
        body = _u8(bed_body)
        si, sp = _idx(sample_idx)
        self.n = n_total if si is None else len(si)
        af = _f64(af)
        m = len(af)
        mi = None if marker_idx is None else np.ascontiguousarray(marker_idx, dtype=np.uint32)
        h = C.c_void_p()
        check(L.saige_grm_create(device, _p(body, _lib.u8p), bytes_per_marker, n_total, sp,
                                 self.n, None if mi is None else _p(mi, _lib.u32p), m,
                                 _p(af, _lib.dp), precision, C.byref(h)))
        self.h = h
        self.m = m
        self.precision = precision

    def mat_vec(self, v):
        v = _f64(v)
        assert v.shape == (self.n,)
        out = np.empty(self.n)
        check(load().saige_grm_mat_vec(self.h, _p(v, _lib.dp), _p(out, _lib.dp)))
        return out

    def mat_mat(self, V):
        """V: (k, n) array of k vectors; returns (k, n)."""
        V = _f64(V)
        k = V.shape[0]
        out = np.empty((k, self.n))
        check(load().saige_grm_mat_mat(self.h, _p(V, _lib.dp), self.n, k, _p(out, _lib.dp),
                                       self.n))
        return out

    def set_marker_chromosomes(self, chrom_codes):
        """Chromosome code (0..C-1) of every GRM marker, for the LOCO products and fits."""
        c = np.ascontiguousarray(chrom_codes, dtype=np.int32)
        check(load().saige_grm_set_marker_chromosomes(self.h, c.ctypes.data_as(
            C.POINTER(C.c_int32))))
        self._n_chrom = int(c.max()) + 1 if c.size else 0

    def mat_mat_loco(self, V, exclude_chrom):
        """GRM product over the markers not on chromosome code ``exclude_chrom``;
        V: (k, n) array of k vectors; returns (k, n)."""
        V = _f64(np.atleast_2d(V))
        k = V.shape[0]
        out = np.empty((k, self.n))
        check(load().saige_grm_mat_mat_loco(self.h, _p(V, _lib.dp), self.n, k,
                                            int(exclude_chrom), _p(out, _lib.dp), self.n))
        return out

    def fit_ai_reml_loco(self, y, x_cm, trait=TRAIT_BINARY, max_iter=30, tol=0.02,
                         pcg_tol=1e-5, pcg_max_iter=500, n_random_vectors=30, seed=12345):
        """run_loco (loco.rs:54-114): one fit_ai_reml per chromosome code, on the GRM
        without that chromosome; the refits share every GRM product.  One dict per code."""
        L = load()
        y = _f64(y)
        x = _f64(x_cm)
        p, n = x.shape
        nc = self._n_chrom
        cfg = _lib.AiRemlConfig(max_iter, tol, pcg_tol, pcg_max_iter, n_random_vectors, seed)
        res = (_lib.AiRemlResult * nc)()
        mu, w, r = np.empty((nc, n)), np.empty((nc, n)), np.empty((nc, n))
        alpha = np.empty((nc, p))
        check(L.saige_fit_ai_reml_loco(self.h, _p(y, _lib.dp), _p(x, _lib.dp), n, p, trait,
                                       C.byref(cfg), nc, res, _p(mu, _lib.dp), _p(w, _lib.dp),
                                       _p(r, _lib.dp), _p(alpha, _lib.dp)))
        return [dict(tau=np.array(res[c].tau[:]), iterations=res[c].iterations,
                     converged=bool(res[c].converged), log_likelihood=res[c].log_likelihood,
                     grm_products=res[c].grm_products, pcg_iterations=res[c].pcg_iterations,
                     mu=mu[c], working_weights=w[c], residuals=r[c], alpha=alpha[c])
                for c in range(nc)]

    def mat_mat_device(self, d_v, ldv, k, d_out, ldo):
        check(load().saige_grm_mat_mat_device(self.h, d_v, ldv, k, d_out, ldo))

    @property
    def stream(self):
        return load().saige_grm_stream(self.h)

    def synchronize(self):
        check(load().saige_grm_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            load().saige_grm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_sigma(grm, w, tau, B, tol=1e-5, max_iter=500):
    """PcgSolver::solve on Sigma = tau0 diag(1/W) + tau1 GRM for k right-hand sides.
    B: (k, n).  Returns (X (k, n), iterations, residuals, converged)."""
    B = _f64(np.atleast_2d(B))
    k, n = B.shape
    w = _f64(w)
    t = _f64(tau)
    X = np.empty((k, n))
    info = (_lib.PcgInfo * k)()
    check(load().saige_pcg_solve_sigma(grm.h, _p(w, _lib.dp), _p(t, _lib.dp), _p(B, _lib.dp), n, k,
                                       tol, max_iter, _p(X, _lib.dp), n, info))
    return (X, np.array([i.iterations for i in info]), np.array([i.residual for i in info]),
            np.array([bool(i.converged) for i in info]))


def fit_ai_reml(grm, y, x_cm, trait=TRAIT_BINARY, max_iter=30, tol=0.02, pcg_tol=1e-5,
                pcg_max_iter=500, n_random_vectors=30, seed=12345):
    """fit_ai_reml (ai_reml.rs:86-427).  x_cm: (p, n), row j = covariate column j."""
    y = _f64(y)
    x = _f64(x_cm)
    p, n = x.shape
    cfg = _lib.AiRemlConfig(max_iter, tol, pcg_tol, pcg_max_iter, n_random_vectors, seed)
    r = _lib.AiRemlResult()
    mu, w, res, eta = (np.empty(n) for _ in range(4))
    alpha = np.empty(p)
    check(load().saige_fit_ai_reml(grm.h, _p(y, _lib.dp), _p(x, _lib.dp), n, p, trait,
                                   C.byref(cfg), C.byref(r), _p(mu, _lib.dp), _p(w, _lib.dp),
                                   _p(res, _lib.dp), _p(eta, _lib.dp), _p(alpha, _lib.dp)))
    return dict(tau=np.array(r.tau[:]), iterations=r.iterations, converged=bool(r.converged),
                log_likelihood=r.log_likelihood, grm_products=r.grm_products,
                pcg_iterations=r.pcg_iterations, pcg_not_converged=r.pcg_not_converged, mu=mu,
                w=w, residuals=res, eta=eta, alpha=alpha)


def xvx_inv_xv(x_cm, w):
    """fit_null.rs:289-307 (flat buffer as the reference fills it)."""
    x = _f64(x_cm)
    p, n = x.shape
    w = _f64(w)
    out = np.empty(p * n)
    check(load().saige_xvx_inv_xv(_p(x, _lib.dp), n, p, _p(w, _lib.dp), _p(out, _lib.dp)))
    return out


def estimate_variance_ratio(grm, dosages, macs, mu, w, tau, x_cm, a_flat, trait=TRAIT_BINARY,
                            n_markers=30, min_mac=20.0, cv_cutoff=1e-3, seed=12345):
    """estimate_variance_ratio (variance_ratio.rs:87-246).  dosages: (m, n)."""
    d = _f64(dosages)
    m, n = d.shape
    x = _f64(x_cm)
    p = x.shape[0]
    macs, mu, w, a, t = _f64(macs), _f64(mu), _f64(w), _f64(a_flat), _f64(tau)
    cfg = _lib.VrConfig(n_markers, min_mac, cv_cutoff, seed)
    vr = C.c_double()
    used = C.c_uint32()
    per = np.empty(max(m, 1))
    check(load().saige_estimate_variance_ratio(grm.h, _p(d, _lib.dp), _p(macs, _lib.dp), m,
                                               _p(mu, _lib.dp), _p(w, _lib.dp), _p(t, _lib.dp),
                                               trait, _p(x, _lib.dp), _p(a, _lib.dp), p,
                                               C.byref(cfg), C.byref(vr), C.byref(used),
                                               _p(per, _lib.dp)))
    return vr.value, used.value, per[:used.value]


def fit_null(bed_body, bytes_per_marker, n_total, m_total, y, x_cm, sample_idx=None,
             trait=TRAIT_BINARY, precision=PREC_FP64, max_iter=30, tol=1e-5, seed=12345,
             n_random_vectors=30, min_maf=0.01, max_missing=0.15, n_markers_vr=30, device=0,
             d_bed=None, use_categorical_vr=False):
    """fit_null::run after phenotype parsing (fit_null.rs:199-394).  With ``d_bed`` (a device
    pointer holding the .bed body) the genotypes are read from HBM instead of host memory."""
    L = load()
    body = _u8(bed_body) if d_bed is None else None
    si, sp = _idx(sample_idx)
    n = n_total if si is None else len(si)
    y = _f64(y)
    x = _f64(x_cm)
    p = x.shape[0]
    cfg = _lib.FitNullConfig(_lib.AiRemlConfig(max_iter, tol, 1e-5, 500, n_random_vectors, seed),
                             min_maf, max_missing, n_markers_vr, precision,
                             int(bool(use_categorical_vr)))
    r = _lib.FitNullResult()
    mu = np.empty(n)
    a = np.empty(p * n)
    cls = np.empty(m_total, np.uint8)
    alpha = np.empty(p)
    pm = np.empty(max(m_total, 1))
    args = (bytes_per_marker, n_total, m_total, sp, n, _p(y, _lib.dp), _p(x, _lib.dp), p, trait,
            C.byref(cfg), C.byref(r), _p(mu, _lib.dp), _p(a, _lib.dp), _p(cls, _lib.u8p),
            _p(alpha, _lib.dp), _p(pm, _lib.dp))
    if d_bed is None:
        check(L.saige_fit_null(device, _p(body, _lib.u8p), *args))
    else:
        check(L.saige_fit_null_device(device, C.c_void_p(d_bed), *args))
    return _fit_null_dict(r, mu, a, cls, alpha, pm)


def _fit_null_dict(r, mu, a, cls, alpha, pm):
    rr = r.reml
    return dict(tau=np.array(rr.tau[:]), iterations=rr.iterations, converged=bool(rr.converged),
                log_likelihood=rr.log_likelihood, grm_products=rr.grm_products,
                pcg_iterations=rr.pcg_iterations, pcg_not_converged=rr.pcg_not_converged,
                vr=r.variance_ratio, vr_markers_used=r.vr_markers_used,
                per_marker_vr=pm[:r.vr_markers_used].copy(), alpha=alpha,
                n_grm_markers=r.n_grm_markers, n_vr_markers=r.n_vr_markers, mu=mu,
                xvx_inv_xv=a, marker_class=cls, seconds_qc=r.seconds_qc,
                seconds_reml=r.seconds_reml, seconds_vr=r.seconds_vr,
                seconds_total=r.seconds_total,
                categorical_vr=[(r.categorical_vr_upper[b], r.categorical_vr[b])
                                for b in range(r.n_categorical_vr)])


def fit_null_multi(bed_body, bytes_per_marker, n_total, m_total, Y, x_cm, traits=None,
                   sample_idx=None, precision=PREC_FP64, max_iter=30, tol=1e-5, seed=12345,
                   n_random_vectors=30, min_maf=0.01, max_missing=0.15, n_markers_vr=30,
                   device=0, d_bed=None, use_categorical_vr=False):
    """Several phenotypes (rows of ``Y``, model sample order) fitted together over one GRM:
    each trait follows its own fit_null, the GRM products of all traits are joined into one
    multi-column product.  Returns one fit_null dict per trait."""
    L = load()
    body = _u8(bed_body) if d_bed is None else None
    si, sp = _idx(sample_idx)
    n = n_total if si is None else len(si)
    Y = np.ascontiguousarray(np.atleast_2d(Y), dtype=np.float64)
    T = Y.shape[0]
    if Y.shape[1] != n:
        raise ValueError("Y must be n_traits x n")
    x = _f64(x_cm)
    p = x.shape[0]
    tt = np.ascontiguousarray([TRAIT_BINARY] * T if traits is None else traits, dtype=np.int32)
    cfg = _lib.FitNullConfig(_lib.AiRemlConfig(max_iter, tol, 1e-5, 500, n_random_vectors, seed),
                             min_maf, max_missing, n_markers_vr, precision,
                             int(bool(use_categorical_vr)))
    res = (_lib.FitNullResult * T)()
    mu = np.empty((T, n))
    a = np.empty((T, p * n))
    cls = np.empty(m_total, np.uint8)
    alpha = np.empty((T, p))
    pm = np.empty((T, max(m_total, 1)))
    args = (bytes_per_marker, n_total, m_total, sp, n, _p(Y, _lib.dp), T, _p(x, _lib.dp), p,
            tt.ctypes.data_as(C.POINTER(C.c_int)), C.byref(cfg), res, _p(mu, _lib.dp),
            _p(a, _lib.dp), _p(cls, _lib.u8p), _p(alpha, _lib.dp), _p(pm, _lib.dp))
    if d_bed is None:
        check(L.saige_fit_null_multi(device, _p(body, _lib.u8p), *args))
    else:
        check(L.saige_fit_null_multi_device(device, C.c_void_p(d_bed), *args))
    return [_fit_null_dict(res[t], mu[t], a[t], cls, alpha[t], pm[t]) for t in range(T)]


def pcg_graphs(enable=None):
    """Query (None) or switch the capture of PCG iterations in HIP graphs; returns the
    previous state."""
    return bool(load().saige_pcg_graphs(-1 if enable is None else int(bool(enable))))


def pcg_graph_counts():
    """(captured, replayed, eager) PCG iterations since the process started."""
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(load().saige_pcg_graph_counts(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def launch_log_enable(on=True, width=0):
    """Start (clearing it) or stop the process-wide launch log of the genotype GEMM launches;
    with width != 0 only launches of that operand width are recorded (diagnostic)."""
    check(load().saige_launch_log_enable(int(bool(on)), int(width)))


def launch_log_summary(cap=64):
    """Launches recorded since the last read, per operand width: {width: (total_ms, launches,
    vectors)}; clears the log (diagnostic)."""
    w = np.zeros(cap, np.uint32)
    ms = np.zeros(cap)
    cnt = np.zeros(cap, np.uint64)
    vec = np.zeros(cap, np.uint64)
    nw = C.c_size_t()
    check(load().saige_launch_log_summary(_p(w, _lib.u32p), _p(ms, _lib.dp),
                                          cnt.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          vec.ctypes.data_as(C.POINTER(C.c_uint64)), cap,
                                          C.byref(nw)))
    return {int(w[i]): (float(ms[i]), int(cnt[i]), int(vec[i])) for i in range(min(nw.value, cap))}


def bgen_dosages(batch, sample_idx=None, device=0):
    """Dosages (nvar x n) of a ``bgen.BgenBatch`` decoded on the device (bgen.rs:373-430)."""
    si, sp = _idx(sample_idx)
    n = batch.n_file if si is None else len(si)
    out = np.empty((batch.n_variants, n))
    check(load().saige_bgen_dosages(device, _p(batch.blocks, _lib.u8p), batch.ld,
                                    batch.n_variants, _p(batch.info, _lib.u32p), batch.n_file,
                                    sp, n, _p(out, _lib.dp)))
    return out


def results_to_dict(arr):
    """Columns of a ctypes array of saige_variant_result as numpy arrays."""
    raw = np.frombuffer(arr, dtype=_RESULT_DTYPE)
    return {f: (raw[f].astype(np.float64) if f == "spa_features" else raw[f].copy())
            for f in _RESULT_DTYPE.names}


def _result_dtype():
    fields = []
    for name, ct in _lib.VariantResult._fields_:
        if name == "spa_features":
            fields.append((name, np.float32, (6,)))
        else:
            fields.append((name, np.dtype(ct)))
    dt = np.dtype({"names": [f[0] for f in fields],
                   "formats": [f[1] if len(f) == 2 else (f[1], f[2]) for f in fields],
                   "offsets": [getattr(_lib.VariantResult, f[0]).offset for f in fields],
                   "itemsize": C.sizeof(_lib.VariantResult)})
    return dt


_RESULT_DTYPE = _result_dtype()


class ScoreTestEngine:
    """ScoreTestEngine (single_variant.rs:77-108) with test_marker over batches."""

    def __init__(self, mu, residuals, tau, x_cm, a_flat, variance_ratio, y=None, use_spa=True,
                 use_fast_spa=True, spa_tol=1e-6, spa_pval_cutoff=0.05, categorical_vr=(),
                 mu2=None, device=0):
        mu = _f64(mu)
        self.n = len(mu)
        mu2 = _f64(mu * (1.0 - mu) if mu2 is None else mu2)
        res = _f64(residuals)
        x = _f64(x_cm)
        p = x.shape[0]
        a = _f64(a_flat)
        yy = None if y is None else _f64(y)
        cu = _f64([c[0] for c in categorical_vr] or [0.0])
        cv = _f64([c[1] for c in categorical_vr] or [0.0])
        h = C.c_void_p()
        check(load().saige_engine_create(device, _p(mu, _lib.dp), _p(mu2, _lib.dp),
                                         _p(res, _lib.dp), tau[0], tau[1], _p(a, _lib.dp),
                                         _p(x, _lib.dp), self.n, p, variance_ratio,
                                         _p(cu, _lib.dp), _p(cv, _lib.dp), len(categorical_vr),
                                         int(use_spa), int(use_fast_spa), spa_tol,
                                         spa_pval_cutoff,
                                         None if yy is None else _p(yy, _lib.dp), C.byref(h)))
        self.h = h

    def test_markers(self, G):
        """G: (nvar, n) f64 dosages (NaN = missing)."""
        G = _f64(np.atleast_2d(G))
        nvar = G.shape[0]
        out = (_lib.VariantResult * nvar)()
        check(load().saige_test_markers_f64(self.h, _p(G, _lib.dp), self.n, nvar, out))
        return results_to_dict(out)

    def test_marker(self, g):
        r = self.test_markers(np.asarray(g)[None, :])
        return {k: v[0] for k, v in r.items()}

    def test_markers_u8(self, Q, scale=1.0 / 127.0):
        Q = _u8(np.atleast_2d(Q))
        nvar, ld = Q.shape
        out = (_lib.VariantResult * nvar)()
        check(load().saige_test_markers_u8(self.h, _p(Q, _lib.u8p), ld, nvar, scale, out))
        return results_to_dict(out)

    def test_markers_plink(self, bed_body, bytes_per_marker, n_total, nvar, sample_idx=None):
        body = _u8(bed_body)
        si, sp = _idx(sample_idx)
        out = (_lib.VariantResult * nvar)()
        check(load().saige_test_markers_plink(self.h, _p(body, _lib.u8p), bytes_per_marker,
                                              n_total, sp, nvar, out))
        return results_to_dict(out)

    def test_markers_bgen(self, batch, sample_idx=None):
        """Variants of a ``bgen.BgenBatch`` (decompressed layout-2 blocks), decoded on the
        device inside the score test; ``sample_idx`` maps model samples to file samples."""
        si, sp = _idx(sample_idx)
        nvar = batch.n_variants
        out = (_lib.VariantResult * nvar)()
        check(load().saige_test_markers_bgen(self.h, _p(batch.blocks, _lib.u8p), batch.ld, nvar,
                                             _p(batch.info, _lib.u32p), batch.n_file, sp, out))
        return results_to_dict(out)

    def test_regions(self, G, region_sizes, weights, rho=None, device_rows=None):
        """BURDEN / SKAT / SKAT-O (region.rs:76-345) for regions of consecutive rows of
        ``G`` ((nvar, n) dosages, NaN = missing) with ``region_sizes`` rows each; returns
        a dict of arrays (pvalue = SKAT-O, burden_pvalue, skat_pvalue, rho, n_markers).
        With ``device_rows=(d_ptr, ld, nvar)`` the rows are read in place from device memory
        (``G`` is then ignored)."""
        if device_rows is None:
            G = _f64(np.atleast_2d(G))
            nvar = G.shape[0]
        else:
            d_ptr, ldg, nvar = device_rows
        sizes = np.asarray(region_sizes, dtype=np.int64)
        starts = np.zeros(len(sizes) + 1, np.uint64)
        starts[1:] = np.cumsum(sizes)
        if int(starts[-1]) != nvar:
            raise ValueError("region sizes must add up to the number of rows of G")
        w = _f64(weights)
        rr = None if rho is None else _f64(rho)
        out = (_lib.RegionResult * max(len(sizes), 1))()
        args = (starts.ctypes.data_as(C.POINTER(C.c_uint64)), len(sizes), _p(w, _lib.dp),
                None if rr is None else _p(rr, _lib.dp), 0 if rr is None else len(rr), out)
        if device_rows is None:
            check(load().saige_region_tests(self.h, _p(G, _lib.dp), self.n, nvar, *args))
        else:
            check(load().saige_region_tests_device(self.h, d_ptr, ldg, nvar, *args))
        raw = np.frombuffer(out, dtype=np.dtype({"names": ["pvalue", "burden_pvalue",
                                                           "skat_pvalue", "rho", "n_markers"],
                                                 "formats": ["f8", "f8", "f8", "f8", "u4"],
                                                 "offsets": [getattr(_lib.RegionResult, f).offset
                                                             for f in ("pvalue", "burden_pvalue",
                                                                       "skat_pvalue", "rho",
                                                                       "n_markers")],
                                                 "itemsize": C.sizeof(_lib.RegionResult)}))
        return {k: raw[k][:len(sizes)].copy() for k in raw.dtype.names}

    def set_root_proposal(self, mlp):
        """Start the SPA Newton iterations from the learned proposal (None: from 0).  Only
        searches the proposal places within 20 of 0 start there, and a search from it that
        fails or ends beyond 20 runs again from 0: every variant keeps the outcome of the
        search from 0 (fast.rs:53-132)."""
        self._mlp = mlp
        check(load().saige_engine_set_root_proposal(self.h, None if mlp is None else mlp.h))

    def set_sums_form(self, form):
        """Kernel of the per-variant sums for uint8 rows: -1 timed choice (default), 0 lanes
        over samples, 1 a lane per variant, 2 int8 matrix cores (exact integer sums against
        a fixed-point expansion of each feature)."""
        check(load().saige_engine_set_sums_form(self.h, int(form)))

    def set_spa_split(self, enable):
        """Carrier-only SPA passes for uint8 variants with few carriers (intercept-only
        engines; True by default).  Returns whether the engine has the zero-byte table they
        need."""
        has = C.c_int(0)
        check(load().saige_engine_set_spa_split(self.h, 1 if enable else 0, C.byref(has)))
        return bool(has.value)

    def copy_tuning_from(self, other):
        """Take other's tuned sums shapes and sums-form and carrier-pass settings (same n and
        covariate count), so that both engines compute a batch the same way."""
        check(load().saige_engine_copy_tuning(self.h, other.h))

    def test_markers_u8_device(self, d_q, ldq, nvar, scale, d_out):
        check(load().saige_test_markers_u8_device(self.h, d_q, ldq, nvar, scale, d_out))

    def test_markers_u8_device_noalloc(self, d_q, ldq, nvar, scale, d_out):
        """test_markers_u8_device with no device allocation (at most 4,096 variants per call;
        work areas in the library's per-GPU scratch array)."""
        check(load().saige_test_markers_u8_device_noalloc(self.h, d_q, ldq, nvar, scale, d_out))

    @property
    def stream(self):
        return load().saige_engine_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            load().saige_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScoreTestPool:
    """Several ScoreTestEngines over one null model, each on its own HIP stream and driven by
    its own host thread, for streams of device-resident uint8 batches.  One engine leaves
    the GPU partly idle: its SPA searches read the number of variants still searching back to
    the host between passes, and their last passes are a few microseconds each.  With two
    engines one batch's HBM-bound sums run beside another batch's SPA passes.

    Batches are dealt round-robin.  The first batch of each new batch size runs alone on
    engine 0, whose tuned sums shape is then copied to the others, so that every batch is
    computed the same way and gives the same bits as on a single engine with that tuning."""

    def __init__(self, *args, engines=2, **kw):
        if engines < 1:
            raise ValueError("engines >= 1")
        self.engines = [ScoreTestEngine(*args, **kw) for _ in range(engines)]
        self.n = self.engines[0].n
        self._sizes = set()

    def _share_tuning(self):
        for e in self.engines[1:]:
            e.copy_tuning_from(self.engines[0])

    def set_sums_form(self, form):
        for e in self.engines:
            e.set_sums_form(form)

    def set_spa_split(self, enable):
        return all([e.set_spa_split(enable) for e in self.engines])

    def test_markers_u8_device(self, batches):
        """batches: (d_q, ldq, nvar, scale, d_out) tuples, each as for
        ScoreTestEngine.test_markers_u8_device.  Returns when every batch's launches are
        issued; each engine's stream then holds its batches' remaining work (synchronize the
        device before reading the results)."""
        import threading
        todo = list(batches)
        new = [b for b in todo if b[2] not in self._sizes]
        for b in new:  # the first batch of a size tunes engine 0's shape for it
            if b[2] not in self._sizes:
                self.engines[0].test_markers_u8_device(*b)
                self._sizes.add(b[2])
                todo.remove(b)
        if new:
            self._share_tuning()
        k = len(self.engines)
        errors = [None] * k

        def work(j):
            try:
                for b in todo[j::k]:
                    self.engines[j].test_markers_u8_device(*b)
            except BaseException as ex:  # re-raised on the calling thread
                errors[j] = ex

        threads = [threading.Thread(target=work, args=(j,)) for j in range(1, k)]
        for t in threads:
            t.start()
        work(0)
        for t in threads:
            t.join()
        for ex in errors:
            if ex is not None:
                raise ex

    def close(self):
        for e in self.engines:
            e.close()


def spa_binary_fast(mu, g, q, qinv, pval_noadj, tol=1e-6, fast=True, device=0):
    """spa_binary_fast (fast.rs:207-242); fast=False gives spa_binary (binary.rs:194-228)."""
    mu, g = _f64(mu), _f64(g)
    r = _lib.SpaResult()
    check(load().saige_spa_binary(device, _p(mu, _lib.dp), _p(g, _lib.dp), len(mu), q, qinv,
                                  pval_noadj, tol, int(fast), C.byref(r)))
    return dict(pvalue=r.pvalue, converged=bool(r.converged), is_spa=bool(r.is_spa),
                iters1=r.iters1, iters2=r.iters2, root1=r.root1, root2=r.root2)


class SpaRootMlp:
    """Learned saddle-point root proposal (no counterpart in saige-rs): a 6 -> 64 -> 64 -> 1
    GELU network trained and run on the GPU; see include/saige_gpu.h."""

    N_PARAMS = 4673

    def __init__(self, seed=0, device=0):
        h = C.c_void_p()
        check(load().saige_mlp_create(device, seed, C.byref(h)))
        self.h = h

    @staticmethod
    def _x(x):
        return np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)

    def params(self):
        p = np.empty(self.N_PARAMS, np.float32)
        nm = np.empty(13, np.float32)
        check(load().saige_mlp_get_params(self.h, _p(p, _lib.fp), _p(nm, _lib.fp)))
        return p, nm

    def set_params(self, p, norm):
        p = np.ascontiguousarray(p, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        check(load().saige_mlp_set_params(self.h, _p(p, _lib.fp), _p(nm, _lib.fp)))

    def predict(self, x):
        x = self._x(x)
        y = np.empty(len(x), np.float32)
        check(load().saige_mlp_predict(self.h, _p(x, _lib.fp), len(x), _p(y, _lib.fp)))
        return y

    def loss_grad(self, x, y):
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float()
        g = np.empty(self.N_PARAMS, np.float32)
        check(load().saige_mlp_loss_grad(self.h, _p(x, _lib.fp), _p(y, _lib.fp), len(x),
                                         C.byref(loss), _p(g, _lib.fp)))
        return loss.value, g

    def train(self, x, y, epochs=20, batch=4096, lr=2e-3, seed=1):
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        hist = np.empty(epochs, np.float32)
        check(load().saige_mlp_train(self.h, _p(x, _lib.fp), _p(y, _lib.fp), len(x), epochs, batch,
                                     lr, seed, _p(hist, _lib.fp)))
        return hist

    def close(self):
        if getattr(self, "h", None):
            load().saige_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spa_training_set(results, tau_e=1.0):
    """(features, targets) of the learned proposal from step-2 results of SPA'd variants:
    both roots of each variant whose root finding converged to a finite saddle point."""
    ok = results["is_spa"].astype(bool) & results["spa_converged"].astype(bool)
    sd = np.sqrt(results["var_t"] / tau_e)
    xs, ys = [], []
    for root, sign in (("spa_root1", 1.0), ("spa_root2", -1.0)):
        t = results[root]
        m = ok & np.isfinite(t) & (sd > 0)
        f = results["spa_features"][m].copy()
        f[:, 0] *= sign
        xs.append(f)
        ys.append(t[m] * sd[m])
    return np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32)
