"""Seeded synthetic cohorts (harness): ctypes wrapper of libsaige_synth.so."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libsaige_synth.so")
_lib = None

dp = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise ImportError(f"{_SO} missing: run python csaige_amd/build.py")
        L = C.CDLL(_SO)
        L.synth_unrelated.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_size_t, C.c_double,
                                      u8p, u8p, dp, dp, dp, dp, dp]
        L.synth_families.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_size_t, C.c_double,
                                     u8p, dp, dp, dp]
        L.synth_null_fit.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
                                     dp, dp, dp]
        L.synth_dosage_u8.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                      C.c_double, C.c_double, u8p, dp]
        L.synth_biobank.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                    dp, C.c_double, C.c_size_t, C.c_double, C.c_double,
                                    C.c_size_t, C.c_double, u8p, dp, dp, dp]
        L.synth_biobank.restype = C.c_int
        L.synth_dosage_mafs.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_double,
                                        C.c_double, dp]
        _lib = L
    return _lib


_gpu = None


def _load_gpu():
    global _gpu
    if _gpu is None:
        so = os.path.join(_HERE, "libsaige_synth_gpu.so")
        if not os.path.exists(so):
            raise ImportError(f"{so} missing: run python csaige_amd/build.py")
        L = C.CDLL(so)
        L.synth_dosage_u8_device.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t,
                                             C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
        L.synth_dosage_u8_device.restype = C.c_int
        _gpu = L
    return _gpu


def dosage_mafs(first, nvar, seed=4, maf=(1e-4, 0.5)):
    """MAFs of step-2 variants [first, first + nvar) (those dosage_u8 draws from)."""
    out = np.empty(nvar)
    _load().synth_dosage_mafs(seed, first, nvar, maf[0], maf[1], _p(out, dp))
    return out


def dosage_u8_device(n, first, nvar, d_mafs, d_q, ld=None, seed=4, stream=None):
    """dosage_u8 generated on the device: rows of variants [first, first + nvar) written to the
    device pointer d_q (nvar x ld bytes) from their MAFs at the device pointer d_mafs
    (dosage_mafs), on `stream` (a hipStream_t; None: the null stream).  Same bytes as
    dosage_u8."""
    ld = n if ld is None else ld
    rc = _load_gpu().synth_dosage_u8_device(seed, n, first, nvar, ld, d_mafs, d_q, stream)
    if rc != 0:
        raise ValueError("synth_dosage_u8_device: bad arguments or launch failure")


def _p(a, t):
    return a.ctypes.data_as(t)


class Cohort(dict):
    """bed: (m, bpm) uint8 .bed body; x: (p, n); y: (n,)."""

    __getattr__ = dict.__getitem__


def unrelated(n=2000, m_grm=5000, m_test=2000, seed=1, maf=(0.01, 0.5), h2=0.3, prevalence=0.10,
              n_cov=2, cov_beta=0.5):
    """Unrelated HWE samples (the 2,000-sample workload)."""
    L = _load()
    bpm = (n + 3) // 4
    bed = np.empty((m_grm, bpm), np.uint8)
    bed_t = np.empty((max(m_test, 1), bpm), np.uint8)
    mg, mt = np.empty(m_grm), np.empty(max(m_test, 1))
    y = np.empty(n)
    p = 1 + n_cov
    x = np.empty((p, n))
    liab = np.empty(n)
    L.synth_unrelated(seed, n, m_grm, m_test, maf[0], maf[1], h2, prevalence, n_cov, cov_beta,
                      _p(bed, u8p), _p(bed_t, u8p) if m_test else None, _p(mg, dp),
                      _p(mt, dp) if m_test else None, _p(y, dp), _p(x, dp), _p(liab, dp))
    return Cohort(n=n, bpm=bpm, bed=bed, bed_test=bed_t[:m_test], maf=mg, maf_test=mt[:m_test],
                  y=y, x=x, liability=liab)


def families(n_fam=10000, kids=3, m=100000, seed=2, maf=(0.01, 0.5), h2=0.4, prevalence=0.05,
             n_cov=2, cov_beta=0.5):
    """Nuclear families (2 founders + kids offspring each)."""
    L = _load()
    n = n_fam * (2 + kids)
    bpm = (n + 3) // 4
    bed = np.empty((m, bpm), np.uint8)
    mf = np.empty(m)
    y = np.empty(n)
    x = np.empty((1 + n_cov, n))
    L.synth_families(seed, n_fam, kids, m, maf[0], maf[1], h2, prevalence, n_cov, cov_beta,
                     _p(bed, u8p), _p(mf, dp), _p(y, dp), _p(x, dp))
    return Cohort(n=n, bpm=bpm, bed=bed, maf=mf, y=y, x=x)


def null_fit(n=400000, p=1, eta_mean=-4.6, eta_sd=0.5, seed=3):
    L = _load()
    mu, y = np.empty(n), np.empty(n)
    x = np.empty((p, n))
    L.synth_null_fit(seed, n, p, eta_mean, eta_sd, _p(mu, dp), _p(y, dp), _p(x, dp))
    return Cohort(n=n, mu=mu, y=y, x=x)


def dosage_u8(n, first, nvar, seed=4, maf=(1e-4, 0.5), ld=None, out=None):
    L = _load()
    ld = n if ld is None else ld
    q = np.empty((nvar, ld), np.uint8) if out is None else out
    mafs = np.empty(nvar)
    L.synth_dosage_u8(seed, n, first, nvar, ld, maf[0], maf[1], _p(q, u8p), _p(mafs, dp))
    return q, mafs


def biobank(n=400000, m=150000, frac_related=0.15, prevalences=(0.005, 0.01, 0.05, 0.20), seed=5,
            maf=(0.01, 0.5), h2=0.4, n_causal=2000, n_cov=2, cov_beta=0.5, bed=None):
    """Several binary phenotypes over one cohort with `frac_related` of the samples in
    half-sibling (2nd-degree) pairs.  ``bed`` may be a preallocated (m, ceil(n/4)) uint8
    array (for instance in pinned memory)."""
    L = _load()
    bpm = (n + 3) // 4
    if bed is None:
        bed = np.empty((m, bpm), np.uint8)
    prev = np.ascontiguousarray(prevalences, dtype=np.float64)
    T = len(prev)
    mf = np.empty(m)
    y = np.empty((T, n))
    x = np.empty((1 + n_cov, n))
    n_pairs = int(round(frac_related * n / 2))
    rc = L.synth_biobank(seed, n, m, n_pairs, T, _p(prev, dp), h2, n_causal, maf[0], maf[1], n_cov,
                         cov_beta, _p(bed, u8p), _p(mf, dp), _p(y, dp), _p(x, dp))
    if rc != 0:
        raise ValueError("synth_biobank: bad arguments")
    return Cohort(n=n, bpm=bpm, bed=bed, maf=mf, y=y, x=x, n_pairs=n_pairs, prevalences=prev)
