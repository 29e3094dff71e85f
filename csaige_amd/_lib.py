"""ctypes declarations of libsaige_gpu.so (include/saige_gpu.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsaige_gpu.so")

SAIGE_OK = 0
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_NOMEM", 4: "ERR_NOT_PD",
          5: "ERR_NO_VR_MARKERS", 6: "ERR_NO_MARKERS"}
PREC_FP64 = 0
PREC_MIXED = 1
PREC_INT8 = 2
PREC_INT8X2 = 3
PREC_F32 = 4
TRAIT_BINARY = 0
TRAIT_QUANTITATIVE = 1
TRAIT_SURVIVAL = 2

dp = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p
fp = C.POINTER(C.c_float)
sz = C.c_size_t


class SaigeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class PcgInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("residual", C.c_double), ("converged", C.c_uint8)]


class AiRemlConfig(C.Structure):
    _fields_ = [("max_iter", C.c_uint32), ("tol", C.c_double), ("pcg_tol", C.c_double),
                ("pcg_max_iter", C.c_uint32), ("n_random_vectors", C.c_uint32),
                ("seed", C.c_uint64)]


class AiRemlResult(C.Structure):
    _fields_ = [("tau", C.c_double * 2), ("iterations", C.c_uint32), ("converged", C.c_uint8),
                ("log_likelihood", C.c_double), ("grm_products", C.c_uint64),
                ("pcg_iterations", C.c_uint64), ("pcg_not_converged", C.c_uint32)]


class VrConfig(C.Structure):
    _fields_ = [("n_markers", C.c_uint32), ("min_mac", C.c_double),
                ("ratio_cv_cutoff", C.c_double), ("seed", C.c_uint64)]


class FitNullConfig(C.Structure):
    _fields_ = [("reml", AiRemlConfig), ("min_maf", C.c_double),
                ("max_missing_rate", C.c_double), ("n_markers_vr", C.c_uint32),
                ("precision", C.c_int), ("use_categorical_vr", C.c_uint32)]


class FitNullResult(C.Structure):
    _fields_ = [("reml", AiRemlResult), ("variance_ratio", C.c_double),
                ("vr_markers_used", C.c_uint32), ("n_grm_markers", C.c_uint32),
                ("n_vr_markers", C.c_uint32), ("seconds_qc", C.c_double),
                ("seconds_reml", C.c_double), ("seconds_vr", C.c_double),
                ("seconds_total", C.c_double), ("n_categorical_vr", C.c_uint32),
                ("categorical_vr_upper", C.c_double * 8), ("categorical_vr", C.c_double * 8)]


class VariantResult(C.Structure):
    _fields_ = [("ac", C.c_double), ("af", C.c_double), ("mac", C.c_double), ("n", C.c_uint64),
                ("score", C.c_double), ("var_t", C.c_double), ("var_t_star", C.c_double),
                ("tstat", C.c_double), ("pvalue", C.c_double), ("pvalue_noadj", C.c_double),
                ("is_spa", C.c_int32), ("spa_converged", C.c_int32), ("beta", C.c_double),
                ("se_beta", C.c_double), ("variance_ratio", C.c_double),
                ("af_cases", C.c_double), ("af_controls", C.c_double),
                ("n_cases", C.c_uint64), ("n_controls", C.c_uint64),
                ("spa_iters1", C.c_int32), ("spa_iters2", C.c_int32),
                ("spa_root1", C.c_double), ("spa_root2", C.c_double),
                ("spa_features", C.c_float * 6)]


class RegionResult(C.Structure):
    _fields_ = [("pvalue", C.c_double), ("burden_pvalue", C.c_double),
                ("skat_pvalue", C.c_double), ("rho", C.c_double), ("n_markers", C.c_uint32)]


class SpaResult(C.Structure):
    _fields_ = [("pvalue", C.c_double), ("converged", C.c_int32), ("is_spa", C.c_int32),
                ("iters1", C.c_int32), ("iters2", C.c_int32), ("root1", C.c_double),
                ("root2", C.c_double)]


# (name, restype, argtypes) for every symbol of include/saige_gpu.h
SIGNATURES = [
    ("saige_last_error", C.c_char_p, []),
    ("saige_release_cached_memory", None, []),
    ("saige_version", C.c_char_p, []),
    ("saige_plink_marker_stats", C.c_int, [C.c_int, u8p, sz, sz, sz, u32p, sz, dp, u32p, dp, dp]),
    ("saige_plink_decode", C.c_int, [C.c_int, u8p, sz, sz, sz, u32p, sz, dp]),
    ("saige_grm_create", C.c_int, [C.c_int, u8p, sz, sz, u32p, sz, u32p, sz, dp, C.c_int,
                                   C.POINTER(vp)]),
    ("saige_grm_create_device", C.c_int, [C.c_int, vp, sz, sz, u32p, sz, u32p, sz, dp, C.c_int,
                                          C.POINTER(vp)]),
    ("saige_grm_destroy", None, [vp]),
    ("saige_grm_n_samples", sz, [vp]),
    ("saige_grm_n_markers", sz, [vp]),
    ("saige_grm_mat_vec", C.c_int, [vp, dp, dp]),
    ("saige_grm_mat_mat", C.c_int, [vp, dp, sz, sz, dp, sz]),
    ("saige_grm_mat_mat_device", C.c_int, [vp, vp, sz, sz, vp, sz]),
    ("saige_grm_stream", vp, [vp]),
    ("saige_grm_synchronize", C.c_int, [vp]),
    ("saige_grm_set_kernel_timing", C.c_int, [vp, C.c_int]),
    ("saige_grm_kernel_timing", C.c_int, [vp, dp, C.POINTER(C.c_uint64)]),
    ("saige_launch_log_enable", C.c_int, [C.c_int, C.c_uint32]),
    ("saige_launch_log_read", C.c_int, [C.c_uint32, dp, C.POINTER(C.c_uint64)]),
    ("saige_launch_log_summary", C.c_int, [C.POINTER(C.c_uint32), dp, C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64), C.c_size_t,
                                          C.POINTER(C.c_size_t)]),
    ("saige_pcg_graphs", C.c_int, [C.c_int]),
    ("saige_pcg_graph_counts", C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64)]),
    ("saige_pcg_solve_sigma", C.c_int, [vp, dp, dp, dp, sz, sz, C.c_double, C.c_uint32, dp, sz,
                                        C.POINTER(PcgInfo)]),
    ("saige_fit_ai_reml", C.c_int, [vp, dp, dp, sz, sz, C.c_int, C.POINTER(AiRemlConfig),
                                    C.POINTER(AiRemlResult), dp, dp, dp, dp, dp]),
    ("saige_grm_set_marker_chromosomes", C.c_int, [vp, C.POINTER(C.c_int32)]),
    ("saige_grm_mat_mat_loco", C.c_int, [vp, dp, sz, sz, C.c_int32, dp, sz]),
    ("saige_fit_ai_reml_loco", C.c_int, [vp, dp, dp, sz, sz, C.c_int, C.POINTER(AiRemlConfig),
                                         sz, C.POINTER(AiRemlResult), dp, dp, dp, dp]),
    ("saige_xvx_inv_xv", C.c_int, [dp, sz, sz, dp, dp]),
    ("saige_estimate_variance_ratio", C.c_int, [vp, dp, dp, sz, dp, dp, dp, C.c_int, dp, dp, sz,
                                                C.POINTER(VrConfig), dp, u32p, dp]),
    ("saige_fit_null", C.c_int, [C.c_int, u8p, sz, sz, sz, u32p, sz, dp, dp, sz, C.c_int,
                                 C.POINTER(FitNullConfig), C.POINTER(FitNullResult), dp, dp,
                                 u8p, dp, dp]),
    ("saige_fit_null_device", C.c_int, [C.c_int, vp, sz, sz, sz, u32p, sz, dp, dp, sz, C.c_int,
                                        C.POINTER(FitNullConfig), C.POINTER(FitNullResult), dp,
                                        dp, u8p, dp, dp]),
    ("saige_fit_null_multi", C.c_int, [C.c_int, u8p, sz, sz, sz, u32p, sz, dp, sz, dp, sz,
                                       C.POINTER(C.c_int), C.POINTER(FitNullConfig),
                                       C.POINTER(FitNullResult), dp, dp, u8p, dp, dp]),
    ("saige_fit_null_multi_device", C.c_int, [C.c_int, vp, sz, sz, sz, u32p, sz, dp, sz, dp, sz,
                                              C.POINTER(C.c_int), C.POINTER(FitNullConfig),
                                              C.POINTER(FitNullResult), dp, dp, u8p, dp, dp]),
    ("saige_engine_create", C.c_int, [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, dp, sz,
                                      sz, C.c_double, dp, dp, sz, C.c_int, C.c_int, C.c_double,
                                      C.c_double, dp, C.POINTER(vp)]),
    ("saige_engine_destroy", None, [vp]),
    ("saige_engine_stream", vp, [vp]),
    ("saige_test_markers_f64", C.c_int, [vp, dp, sz, sz, C.POINTER(VariantResult)]),
    ("saige_test_markers_u8", C.c_int, [vp, u8p, sz, sz, C.c_double, C.POINTER(VariantResult)]),
    ("saige_test_markers_u8_device", C.c_int, [vp, vp, sz, sz, C.c_double, vp]),
    ("saige_test_markers_u8_device_noalloc", C.c_int, [vp, vp, sz, sz, C.c_double, vp]),
    ("saige_test_markers_plink", C.c_int, [vp, u8p, sz, sz, u32p, sz,
                                           C.POINTER(VariantResult)]),
    ("saige_bgen_dosages", C.c_int, [C.c_int, u8p, sz, sz, u32p, C.c_uint32, u32p, sz, dp]),
    ("saige_test_markers_bgen", C.c_int, [vp, u8p, sz, sz, u32p, C.c_uint32, u32p,
                                          C.POINTER(VariantResult)]),
    ("saige_region_tests", C.c_int, [vp, dp, sz, sz, C.POINTER(C.c_uint64), sz, dp, dp, sz,
                                     C.POINTER(RegionResult)]),
    ("saige_region_tests_device", C.c_int, [vp, vp, sz, sz, C.POINTER(C.c_uint64), sz, dp, dp,
                                            sz, C.POINTER(RegionResult)]),
    ("saige_chi2_sf", C.c_double, [C.c_double, C.c_double]),
    ("saige_mlp_create", C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
    ("saige_mlp_destroy", None, [vp]),
    ("saige_mlp_set_params", C.c_int, [vp, fp, fp]),
    ("saige_mlp_get_params", C.c_int, [vp, fp, fp]),
    ("saige_mlp_predict", C.c_int, [vp, fp, sz, fp]),
    ("saige_mlp_loss_grad", C.c_int, [vp, fp, fp, sz, fp, fp]),
    ("saige_mlp_train", C.c_int, [vp, fp, fp, sz, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                  fp]),
    ("saige_engine_set_root_proposal", C.c_int, [vp, vp]),
    ("saige_engine_set_sums_form", C.c_int, [vp, C.c_int]),
    ("saige_engine_set_spa_split", C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    ("saige_engine_copy_tuning", C.c_int, [vp, vp]),
    ("saige_spa_binary", C.c_int, [C.c_int, dp, dp, sz, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.POINTER(SpaResult)]),
]

_lib = None


def load():
    """Load libsaige_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python csaige_amd/build.py` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != SAIGE_OK:
        msg = load().saige_last_error().decode(errors="replace")
        raise SaigeError(status, msg)
