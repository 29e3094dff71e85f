/*
 * saige_gpu.h — C ABI of the MI355X (gfx950) implementation of saige-rs's two
 * data-parallel hot paths: the fit-null AI-REML loop with the on-the-fly GRM product,
 * and the step-2 single-variant score test with variance ratio and saddle-point
 * approximation.
 *
 * Conventions (mirroring the saige-rs Rust API they replace):
 *   - Inputs are borrowed, caller-owned buffers (Rust `&[f64]` slices); outputs are
 *     caller-allocated buffers the library fills (Rust returns owned `Vec`s).
 *   - Matrices are column-major like saige-linalg's DenseMatrix (dense.rs:10-17).
 *   - Every entry point returns an int status (saige_status).  Nothing panics or throws
 *     across the ABI; the message of the last failure on the calling thread is
 *     saige_last_error().  PCG non-convergence is a flag in the result, not an error,
 *     as in PcgSolver::solve (decomposition.rs:291-297).
 *   - Host-pointer entry points copy in and out; `_device` entry points take device
 *     pointers and run on the handle's HIP stream.
 *   - A handle may be used from several threads; calls on one handle are serialized
 *     internally (the Send + Sync closure contract of ai_reml.rs:93-95).
 *   - Reference citations are relative to the saige-rs workspace (saige-rs/...).
 */
#ifndef SAIGE_GPU_H
#define SAIGE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SAIGE_OK = 0,
  SAIGE_ERR_ARG = 1,            /* bad dimension / null pointer (reference: assert_eq! panics) */
  SAIGE_ERR_DEVICE = 2,         /* HIP runtime error or no usable GPU */
  SAIGE_ERR_NOMEM = 3,          /* device allocation failed */
  SAIGE_ERR_NOT_PD = 4,         /* LinalgError::NotPositiveDefinite (decomposition.rs:17-18) */
  SAIGE_ERR_NO_VR_MARKERS = 5,  /* "No markers available for variance ratio" (variance_ratio.rs:214-216) */
  SAIGE_ERR_NO_MARKERS = 6      /* no marker passes QC */
} saige_status;

enum { SAIGE_PREC_FP64 = 0, SAIGE_PREC_MIXED = 1, SAIGE_PREC_INT8 = 2, SAIGE_PREC_INT8X2 = 3,
       SAIGE_PREC_F32 = 4 };
enum { SAIGE_TRAIT_BINARY = 0, SAIGE_TRAIT_QUANTITATIVE = 1, SAIGE_TRAIT_SURVIVAL = 2 };

/* Message of the last failed call on this thread ("" if none). */
const char* saige_last_error(void);

/* Device memory is allocated through a per-device cache: buffers a handle or a call frees
 * are kept (up to 32 GiB) and reused by later calls of a similar size, and the cache is
 * emptied whenever an allocation would otherwise fail.  This releases every cached block
 * (no counterpart in saige-rs). */
void saige_release_cached_memory(void);
/* Library version string. */
const char* saige_version(void);

/* ------------------------------------------------------------------------------
 * Genotype decode and allele counts.
 *   Replaces PlinkReader::read_marker (saige-geno/src/plink.rs:237-267, decode_genotype
 *   :170-179, read_marker_raw :182-194) and MarkerData::compute_af (traits.rs:42-54).
 *   bed_body: SNP-major .bed body without the 3 magic bytes; marker j starts at
 *   j * bytes_per_marker; sample s at bits 2*(s%4) of byte s/4; codes 00->2, 01->NaN,
 *   10->1, 11->0 (dosage of bim allele 1).  sample_idx (length n, or NULL for all
 *   n_total samples in file order) is the set_sample_subset index list (plink.rs:223-235).
 * ---------------------------------------------------------------------------- */
int saige_plink_marker_stats(int device, const uint8_t* bed_body, size_t bytes_per_marker,
                             size_t n_total, size_t m, const uint32_t* sample_idx, size_t n,
                             double* ac, uint32_t* n_valid, double* af, double* mac);
/* Dosages of m markers (n x m col-major, NaN = missing). */
int saige_plink_decode(int device, const uint8_t* bed_body, size_t bytes_per_marker,
                       size_t n_total, size_t m, const uint32_t* sample_idx, size_t n,
                       double* dosages);

/* ------------------------------------------------------------------------------
 * On-the-fly GRM operator.
 *   Replaces OnTheFlyGrm::new (saige-core/src/glmm/pcg.rs:101-122) and
 *   OnTheFlyGrm::mat_vec (pcg.rs:128-180): GRM * v = (1/M) * sum_m z_m (z_m' v) with
 *   z_im = (g_im - 2 af_m) / sqrt(2 af_m (1 - af_m)), missing -> 0, monomorphic -> 0.
 *   Takes the packed 2-bit genotypes (resident in HBM) instead of f64 dosages; af is
 *   compute_af of each selected marker over the n samples.
 *   precision SAIGE_PREC_FP64: f64 standardized values, f64 accumulation.
 *   precision SAIGE_PREC_MIXED: exact genotypes on bf16 MFMA with the standardization
 *   factored out, v split into bf16 hi + lo, fp32 accumulation (error stated per output).
 *   precision SAIGE_PREC_INT8: exact genotypes on int8 MFMA with the standardization
 *   factored out, the centred v split into three int8 slices with per-column scales
 *   (|v - split| <= 6.1e-8 max|v - mean|), exact int32 accumulation; n, M <= 8,000,000.
 *   precision SAIGE_PREC_INT8X2: as SAIGE_PREC_INT8 with two int8 slices
 *   (|v - split| <= 1.55e-5 max|v - mean|): 2/3 of the MFMA work.  The fits take three
 *   slices for the narrow solves and, once the relative tau change is below 1e-3, for the
 *   y~ and X columns of the main block solves (the probe columns keep two).
 *   The factored modes correct markers with missing calls by sparse per-marker and
 *   per-sample lists (mean imputation as pcg.rs:108-113).
 *   precision SAIGE_PREC_F32: saige-rs's own arithmetic (pcg.rs:128-180 at one rayon
 *   thread): f32 standardized values, f32 sums in sample order per marker and in marker
 *   order per sample, no contraction; each product bit-identical to the reference's.
 *   No LOCO products in this mode.
 * ---------------------------------------------------------------------------- */
typedef struct saige_grm saige_grm;

int saige_grm_create(int device, const uint8_t* bed_body, size_t bytes_per_marker,
                     size_t n_total, const uint32_t* sample_idx, size_t n,
                     const uint32_t* marker_idx, size_t m, const double* af, int precision,
                     saige_grm** out);
/* Same, from a .bed body already resident in device memory (d_bed_body). */
int saige_grm_create_device(int device, const uint8_t* d_bed_body, size_t bytes_per_marker,
                            size_t n_total, const uint32_t* sample_idx, size_t n,
                            const uint32_t* marker_idx, size_t m, const double* af,
                            int precision, saige_grm** out);
void saige_grm_destroy(saige_grm* h);
size_t saige_grm_n_samples(const saige_grm* h);
size_t saige_grm_n_markers(const saige_grm* h);
/* OnTheFlyGrm::mat_vec (pcg.rs:128-180): out = GRM * v, host vectors of length n. */
int saige_grm_mat_vec(saige_grm* h, const double* v, double* out);
/* k products at once: out[:, c] = GRM * v[:, c] (col-major, leading dimensions ldv, ldo). */
int saige_grm_mat_mat(saige_grm* h, const double* v, size_t ldv, size_t k, double* out,
                      size_t ldo);
/* Same on device pointers, enqueued on the handle's stream. */
int saige_grm_mat_mat_device(saige_grm* h, const double* d_v, size_t ldv, size_t k,
                             double* d_out, size_t ldo);
/* The handle's HIP stream (hipStream_t) and a device synchronize helper. */
void* saige_grm_stream(saige_grm* h);
int saige_grm_synchronize(saige_grm* h);
/* Kernel timing (diagnostic, no counterpart in saige-rs): while enabled, the handle records
 * HIP events on its stream around every launch of its genotype GEMM kernels (two per
 * product: the marker-side and the sample-side pass, both precisions).
 * saige_grm_kernel_timing waits for the last recorded launch and returns the summed
 * launch time and the number of launches since the previous read (or enable), then resets.
 * At most 4096 launches are recorded between reads. */
int saige_grm_set_kernel_timing(saige_grm* h, int enable);
int saige_grm_kernel_timing(saige_grm* h, double* total_ms, uint64_t* launches);
/* Launch log (diagnostic, no counterpart in saige-rs): while enabled, every genotype GEMM
 * launch of every GRM handle in the process, including the handles that the fit-null, LOCO
 * and several-phenotype entry points create internally, is timed with HIP events on its
 * stream and filed under its operand width: the padded number of dense columns it
 * multiplies (int8: 16 * ceil(S * k / 16) slice columns for S operand slices; bf16:
 * 16 * ceil(k / 16); fp64: the column block of the launch).  A handle whose own kernel
 * timing is on records there instead.  saige_launch_log_enable clears the log and records
 * only the launches of the given width from then on (0: every width; only recorded launches
 * pay for the event pair).
 * saige_launch_log_read waits for the recorded launches, returns the summed time and the
 * count of the launches of one width (0: every width) and clears the log; call it when no
 * launch is being issued on another thread.  At most 65536 launches are kept between
 * reads. */
int saige_launch_log_enable(int enable, uint32_t width);
int saige_launch_log_read(uint32_t width, double* total_ms, uint64_t* launches);
/* Every width at once (diagnostic, no counterpart in saige-rs): for each operand width
 * recorded since the last read, in increasing order, the summed launch time, the launch
 * count and the summed number of vectors (the columns of the product before slicing and
 * padding) of its launches; at most cap widths are written, *n_widths is the number recorded.
 * Clears the log like saige_launch_log_read. */
int saige_launch_log_summary(uint32_t* widths, double* total_ms, uint64_t* launches,
                             uint64_t* vectors, size_t cap, size_t* n_widths);

/* ------------------------------------------------------------------------------
 * Preconditioned conjugate gradient on Sigma = tau_e * diag(1/W) + tau_g * GRM.
 *   Replaces PcgSolver::solve (saige-linalg/src/decomposition.rs:212-298) called with
 *   the Sigma / diagonal-preconditioner closures of ai_reml.rs:142-164 (and
 *   fit_null.rs:320-341).  k right-hand sides are solved together; each keeps its own
 *   alpha, beta and stopping test, so every column is iterate-identical to a separate
 *   PcgSolver::solve (x0 = 0, stop when ||r|| < tol * max(||b||, 1), breakdown when
 *   |p'Ap| < 1e-30).
 * ---------------------------------------------------------------------------- */
typedef struct {
  uint32_t iterations;   /* PcgResult::iterations */
  double residual;       /* PcgResult::residual   */
  uint8_t converged;     /* PcgResult::converged  */
} saige_pcg_info;

int saige_pcg_solve_sigma(saige_grm* h, const double* w, const double tau[2], const double* b,
                          size_t ldb, size_t k, double tol, uint32_t max_iter, double* x,
                          size_t ldx, saige_pcg_info* info);
/* PCG iteration graphs (no counterpart in saige-rs): the launches of one PCG iteration at one
 * active width (GRM product of the active columns, Sigma, p'Ap, update, beta) are captured in
 * a HIP graph the second time that iteration shape is met and replayed from then on, while
 * no device workspace has moved; the host reads the active count between graph launches.
 * Single-trait solves only (the joined products of several traits or LOCO refits run
 * eagerly), and not while kernel timing or the launch log is on.  The results are those of
 * the uncaptured launches bit for bit.  enable: 1 on (the default), 0 off, -1 query; returns
 * the previous state. */
int saige_pcg_graphs(int enable);
/* PCG iterations run since the process started: captured into a new graph, replayed from
 * one, and launched without a graph (diagnostic). */
int saige_pcg_graph_counts(uint64_t* captured, uint64_t* replayed, uint64_t* eager);

/* ------------------------------------------------------------------------------
 * AI-REML null-model fit.
 *   Replaces fit_ai_reml (saige-core/src/glmm/ai_reml.rs:86-427) with AiRemlConfig
 *   (ai_reml.rs:22-51) and AiRemlResult (ai_reml.rs:55-74).  The GRM closure is the
 *   handle.  Hutchinson probes are the ChaCha8Rng(seed) Rademacher draws of
 *   ai_reml.rs:121-130.  Outputs mu, w, residuals, eta (length n) and alpha (length p)
 *   may be NULL.
 * ---------------------------------------------------------------------------- */
typedef struct {
  uint32_t max_iter;          /* default 30 */
  double tol;                 /* default 0.02 (the CLI passes 1e-5, fit_null.rs:78-80) */
  double pcg_tol;             /* default 1e-5 */
  uint32_t pcg_max_iter;      /* default 500 */
  uint32_t n_random_vectors;  /* default 30 */
  uint64_t seed;              /* default 12345 */
} saige_ai_reml_config;

typedef struct {
  double tau[2];
  uint32_t iterations;
  uint8_t converged;
  double log_likelihood;
  /* diagnostics (no reference counterpart) */
  uint64_t grm_products;      /* GRM column-products performed */
  uint64_t pcg_iterations;    /* summed over right-hand sides */
  uint32_t pcg_not_converged; /* solves that hit max_iter or broke down */
} saige_ai_reml_result;

int saige_fit_ai_reml(saige_grm* h, const double* y, const double* x, size_t n, size_t p,
                      int trait_type, const saige_ai_reml_config* cfg,
                      saige_ai_reml_result* result, double* mu, double* w, double* residuals,
                      double* eta, double* alpha);

/* Leave-one-chromosome-out (saige-core/src/glmm/loco.rs:54-114).
 *   saige_grm_set_marker_chromosomes: chromosome code per GRM marker (handle order), codes
 *     0..C-1 each used by at least one marker (NULL clears).  The host maps chromosome names
 *     to codes in sorted order, the order run_loco fits them (loco.rs:64-71).
 *   saige_grm_mat_mat_loco: the GRM product over the markers not on chromosome
 *     exclude_chrom (OnTheFlyGrm::new on the kept markers, loco.rs:78-96): the same resident
 *     genotypes with that chromosome's rows of Z^T v set to 0 and 1/M_c for 1/M.
 *   saige_fit_ai_reml_loco: fit_ai_reml once per chromosome code on the GRM without it
 *     (loco.rs:98-111); the C refits run together and share every product.  Outputs per
 *     code c: results[c], and (each may be NULL) mu, w, residuals (n x C) and alpha (p x C). */
int saige_grm_set_marker_chromosomes(saige_grm* h, const int32_t* chrom);
int saige_grm_mat_mat_loco(saige_grm* h, const double* v, size_t ldv, size_t k,
                           int32_t exclude_chrom, double* out, size_t ldo);
int saige_fit_ai_reml_loco(saige_grm* h, const double* y, const double* x, size_t n, size_t p,
                           int trait_type, const saige_ai_reml_config* cfg, size_t n_chrom,
                           saige_ai_reml_result* results, double* mu, double* w,
                           double* residuals, double* alpha);

/* ------------------------------------------------------------------------------
 * (X'WX)^{-1} X'W as produced by fit_null.rs:289-307 (a p*n flat buffer filled
 * element (j, i) -> [j*n + i], which every consumer reads as col-major p x n,
 * fit_null.rs:307 and assoc_test.rs:125; for p > 1 this is the reference's layout,
 * reproduced as is).
 * ---------------------------------------------------------------------------- */
int saige_xvx_inv_xv(const double* x, size_t n, size_t p, const double* w, double* out);

/* ------------------------------------------------------------------------------
 * Variance ratio.
 *   Replaces estimate_variance_ratio (saige-core/src/glmm/variance_ratio.rs:87-246) with
 *   VarianceRatioConfig (:30-59, overall ratio; categorical bins not built) and the
 *   Sigma^{-1} PCG closure of fit_null.rs:320-343.  dosages: n x m col-major f64.
 * ---------------------------------------------------------------------------- */
typedef struct {
  uint32_t n_markers;        /* default 30 */
  double min_mac;            /* default 20 */
  double ratio_cv_cutoff;    /* default 1e-3 */
  uint64_t seed;             /* default 12345 */
} saige_vr_config;

int saige_estimate_variance_ratio(saige_grm* h, const double* dosages, const double* mac,
                                  size_t m, const double* mu, const double* w,
                                  const double tau[2], int trait_type, const double* x,
                                  const double* xvx_inv_xv, size_t p,
                                  const saige_vr_config* cfg, double* variance_ratio,
                                  uint32_t* n_used, double* per_marker_vr);

/* ------------------------------------------------------------------------------
 * Whole step 1: saige-cli fit_null::run (saige-cli/src/commands/fit_null.rs:199-394)
 * after phenotype parsing: marker QC (min_maf, max_missing), the ChaCha8 VR/GRM split
 * (:230-259), GRM build, AI-REML, XVX^{-1}X'V and the variance ratio.  y and x are in
 * model sample order (the sample_idx order).  Outputs (each may be NULL): mu (n),
 * xvx_inv_xv (p*n, NullModel::xvx_inv_xv_flat), marker_class (m_total: 0 failed QC, 1 GRM,
 * 2 VR pool), alpha (p fixed effects, AiRemlResult::alpha) and per_marker_vr (room for
 * m_total values; the first result->vr_markers_used are VarianceRatioResult::per_marker_vr)
 * -- with tau and the variance ratio these fill NullModel (null_model.rs:17-60).
 * ---------------------------------------------------------------------------- */
typedef struct {
  saige_ai_reml_config reml;
  double min_maf;            /* 0.01 */
  double max_missing_rate;   /* 0.15 */
  uint32_t n_markers_vr;     /* 30 */
  int precision;             /* SAIGE_PREC_FP64 | _MIXED | _INT8 | _INT8X2 | _F32 */
  uint32_t use_categorical_vr; /* --use-categorical-vr (fit_null.rs:94-96, 348): MAC-binned ratios */
} saige_fit_null_config;

typedef struct {
  saige_ai_reml_result reml;
  double variance_ratio;
  uint32_t vr_markers_used;
  uint32_t n_grm_markers;
  uint32_t n_vr_markers;
  double seconds_qc, seconds_reml, seconds_vr, seconds_total;
  /* VarianceRatioResult::categorical_vr (variance_ratio.rs:221-230, 263-287): the bins of
   * the reference's default MAC boundaries that hold a used marker, as (upper MAC bound,
   * mean ratio) pairs in bin order; n_categorical_vr is 0 unless use_categorical_vr. */
  uint32_t n_categorical_vr;
  double categorical_vr_upper[8];
  double categorical_vr[8];
} saige_fit_null_result;

int saige_fit_null(int device, const uint8_t* bed_body, size_t bytes_per_marker, size_t n_total,
                   size_t m_total, const uint32_t* sample_idx, size_t n, const double* y,
                   const double* x, size_t p, int trait_type, const saige_fit_null_config* cfg,
                   saige_fit_null_result* result, double* mu, double* xvx_inv_xv,
                   uint8_t* marker_class, double* alpha, double* per_marker_vr);

/* Same, with the .bed body already resident in device memory. */
int saige_fit_null_device(int device, const uint8_t* d_bed_body, size_t bytes_per_marker,
                          size_t n_total, size_t m_total, const uint32_t* sample_idx, size_t n,
                          const double* y, const double* x, size_t p, int trait_type,
                          const saige_fit_null_config* cfg, saige_fit_null_result* result,
                          double* mu, double* xvx_inv_xv, uint8_t* marker_class, double* alpha,
                          double* per_marker_vr);

/* Several phenotypes over the same samples, covariates and GRM: n_traits runs of
 * fit_null::run (fit_null.rs:199-394, one per trait, each with its own W, tau, PCG
 * iterations and variance ratio) fitted together.  Each trait runs on its own host thread;
 * whenever every trait still running has a GRM product pending, the products are joined
 * into one multi-column product (its own Sigma epilogue per trait), so each trait follows
 * its single-trait iteration while the genotypes are streamed once per joined product.
 * y: n x n_traits col-major; trait_types[n_traits]; results[n_traits];
 * mu: n x n_traits; xvx_inv_xv: (p*n) x n_traits; marker_class: m_total (shared);
 * alpha: p x n_traits; per_marker_vr: m_total x n_traits.
 * No counterpart in saige-rs beyond calling fit_null once per phenotype. */
int saige_fit_null_multi(int device, const uint8_t* bed_body, size_t bytes_per_marker,
                         size_t n_total, size_t m_total, const uint32_t* sample_idx, size_t n,
                         const double* y, size_t n_traits, const double* x, size_t p,
                         const int* trait_types, const saige_fit_null_config* cfg,
                         saige_fit_null_result* results, double* mu, double* xvx_inv_xv,
                         uint8_t* marker_class, double* alpha, double* per_marker_vr);
int saige_fit_null_multi_device(int device, const uint8_t* d_bed_body, size_t bytes_per_marker,
                                size_t n_total, size_t m_total, const uint32_t* sample_idx,
                                size_t n, const double* y, size_t n_traits, const double* x,
                                size_t p, const int* trait_types,
                                const saige_fit_null_config* cfg, saige_fit_null_result* results,
                                double* mu, double* xvx_inv_xv, uint8_t* marker_class,
                                double* alpha, double* per_marker_vr);

/* ------------------------------------------------------------------------------
 * Step 2: score test engine.
 *   Replaces ScoreTestEngine (saige-core/src/score_test/single_variant.rs:77-108),
 *   test_marker (:112-257) with compute_case_control_af (:260-291), and the fast SPA
 *   spa_binary_fast (saige-core/src/spa/fast.rs:207-242; spa_binary when use_fast_spa
 *   is 0 gives the same result: fast.rs:192 sets apart samples with |g~| <= 1e-10,
 *   and the covariate-adjusted g~ of single_variant.rs:200 is practically never that
 *   small; the kernel applies the test literally).  x: n x p col-major; xvx_inv_xv: the p*n
 *   flat buffer of the null model (NullModel::xvx_inv_xv_flat, null_model.rs:46-47);
 *   categorical ratios are (upper MAC bound, ratio) pairs (variance_ratio.rs:289-301);
 *   y may be NULL (no case/control AF).
 * ---------------------------------------------------------------------------- */
typedef struct saige_engine saige_engine;

typedef struct {
  double ac, af, mac;
  uint64_t n;
  double score, var_t, var_t_star, tstat, pvalue, pvalue_noadj;
  int32_t is_spa, spa_converged;
  double beta, se_beta, variance_ratio, af_cases, af_controls;
  uint64_t n_cases, n_controls;
  int32_t spa_iters1, spa_iters2;  /* Newton iterations of the two roots (diagnostic) */
  double spa_root1, spa_root2;     /* saddle points of q and qinv (diagnostic)          */
  float spa_features[6];           /* inputs of the learned root proposal (diagnostic)  */
} saige_variant_result;            /* numeric fields of SingleVariantResult (:23-74) */

/* Region tests (saige-core/src/score_test/region.rs:76-345): BURDEN, SKAT and SKAT-O with
 * the Satterthwaite moment match of the reference, for regions of consecutive dosage rows.
 *   g: nvar x ldg (row per variant, NaN = missing, which region tests read as 0,
 *      region.rs:86-94); region_start[n_regions + 1]: first row of each region, the last
 *      entry = nvar; weights[nvar]: per-variant weights (beta_weights(MAF, 1, 25) in
 *      saige test, assoc_test.rs:299); rho[n_rho]: the SKAT-O grid (NULL: the reference's
 *      default 0, 0.1, ..., 1).  pvalue is SKAT-O's (the smallest over rho, whose rho is
 *      reported).  p-values are 1 - ChiSquared(df).cdf(x) as the reference computes them,
 *      the CDF through the regularized incomplete gamma function. */
typedef struct {
  double pvalue, burden_pvalue, skat_pvalue;
  double rho;
  uint32_t n_markers;
} saige_region_result;

int saige_engine_create(int device, const double* mu, const double* mu2,
                        const double* residuals, double tau_e, double tau_g,
                        const double* xvx_inv_xv, const double* x, size_t n, size_t p,
                        double variance_ratio, const double* cat_upper, const double* cat_vr,
                        size_t n_cat, int use_spa, int use_fast_spa, double spa_tol,
                        double spa_pval_cutoff, const double* y, saige_engine** out);
void saige_engine_destroy(saige_engine* e);
void* saige_engine_stream(saige_engine* e);

/* test_marker over nvar dosage vectors g (n x nvar col-major f64, NaN = missing). */
int saige_test_markers_f64(saige_engine* e, const double* g, size_t ldg, size_t nvar,
                           saige_variant_result* out);
/* uint8-quantized dosages (dosage = q * scale), variant-major rows of ldq bytes. */
int saige_test_markers_u8(saige_engine* e, const uint8_t* q, size_t ldq, size_t nvar,
                          double scale, saige_variant_result* out);
int saige_test_markers_u8_device(saige_engine* e, const uint8_t* d_q, size_t ldq, size_t nvar,
                                 double scale, saige_variant_result* d_out);
/* saige_test_markers_u8_device without any device allocation.  No counterpart in saige-rs:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The batch's work areas (sums partials,
 * gated-variant list, batched SPA state) come from a file-scope device array of 48 MiB,
 * one per GPU, held until the process ends; at most 4,096 variants per call (more is
 * SAIGE_ERR_ARG), the sums kernel's shape is the one the engine has tuned for that batch
 * size when it has at most 64 sample splits, else the default (no shape search).  Calls on
 * one GPU share its array and keep nothing between calls; calls that overlap on one GPU
 * (engines used from several host threads) race. */
int saige_test_markers_u8_device_noalloc(saige_engine* e, const uint8_t* d_q, size_t ldq,
                                         size_t nvar, double scale,
                                         saige_variant_result* d_out);
/* Hard calls straight from a .bed body (PlinkReader::read_marker + test_marker). */
int saige_test_markers_plink(saige_engine* e, const uint8_t* bed_body, size_t bytes_per_marker,
                             size_t n_total, const uint32_t* sample_idx, size_t nvar,
                             saige_variant_result* out);

/* BGEN v1.2 layout-2 genotype data (BgenReader::read_marker, saige-geno/src/bgen.rs:314-437
 * with the subset of :466-480), decoded on the device inside the score test.
 *   blocks: nvar decompressed layout-2 probability blocks, one every ld bytes, zero-filled
 *           past each block's data (read_bits reads 0 there, bgen.rs:439-452); every
 *           block's N (its first u32) must equal n_file.
 *   info[v]: bits per probability (1..32; 0 is the reference's error) | 0x100 when the
 *           variant is diploid, biallelic and unphased -- the case whose dosage
 *           P(AB) + 2 (1 - P(AA) - P(AB)) is decoded; other variants read 0, and a missing
 *           sample (bit 7 of its ploidy byte) NaN, as bgen.rs:408-430.
 *   sample_idx: model sample -> sample of the file (NULL: all n_file samples, in order).
 * saige_bgen_dosages writes the decoded dosages (n x nvar, column per variant). */
int saige_bgen_dosages(int device, const uint8_t* blocks, size_t ld, size_t nvar,
                       const uint32_t* info, uint32_t n_file, const uint32_t* sample_idx,
                       size_t n, double* dosages);
int saige_test_markers_bgen(saige_engine* e, const uint8_t* blocks, size_t ld, size_t nvar,
                            const uint32_t* info, uint32_t n_file, const uint32_t* sample_idx,
                            saige_variant_result* out);
int saige_region_tests(saige_engine* e, const double* g, size_t ldg, size_t nvar,
                       const uint64_t* region_start, size_t n_regions, const double* weights,
                       const double* rho, size_t n_rho, saige_region_result* out);
/* The same with the dosage rows resident on the engine's device (d_g: nvar rows of ldg
 * doubles in device memory, read in place; weights, region_start and out stay on the host). */
int saige_region_tests_device(saige_engine* e, const double* d_g, size_t ldg, size_t nvar,
                              const uint64_t* region_start, size_t n_regions,
                              const double* weights, const double* rho, size_t n_rho,
                              saige_region_result* out);
/* 1 - ChiSquared::new(df).cdf(x) as the region tests evaluate it (host arithmetic). */
double saige_chi2_sf(double df, double x);

/* ------------------------------------------------------------------------------
 * Learned saddle-point root proposal (no counterpart in saige-rs).
 *   A 6 -> 64 -> 64 -> 1 GELU network proposes the Newton starting points of the two SPA
 *   roots from per-variant features (saige_variant_result::spa_features); the
 *   safeguarded Newton of find_root_k1_fast (fast.rs:53-132) still refines every root to
 *   the engine's tolerance.  Parameters: 4,673 fp32 master weights (layout in
 *   csaige_amd/csrc/mlp_common.hpp), inference with bf16-rounded weights and fp32
 *   accumulation.  norm[13] = feature means (6), inverse standard deviations (6), y_scale.
 *   Inputs x are n x 6 row-major raw features; targets y are t * sqrt(varT / tau_e),
 *   fitted through asinh(y) / y_scale (predict returns y).
 * ---------------------------------------------------------------------------- */
typedef struct saige_mlp saige_mlp;

int saige_mlp_create(int device, uint64_t seed, saige_mlp** out);
void saige_mlp_destroy(saige_mlp* h);
int saige_mlp_set_params(saige_mlp* h, const float* params, const float* norm);
int saige_mlp_get_params(saige_mlp* h, float* params, float* norm);
/* y = network output in target units (bf16 weights, fp32 accumulation). */
int saige_mlp_predict(saige_mlp* h, const float* x, size_t n, float* y);
/* Mean squared error on asinh(y) / y_scale and its gradient w.r.t. the fp32 parameters. */
int saige_mlp_loss_grad(saige_mlp* h, const float* x, const float* y, size_t n, float* loss,
                        float* grad);
/* Adam training on the GPU (standardization fitted to x, y); loss_history[epochs]. */
int saige_mlp_train(saige_mlp* h, const float* x, const float* y, size_t n, uint32_t epochs,
                    uint32_t batch, float lr, uint64_t seed, float* loss_history);
/* Use the proposal as the SPA starting point of an engine (NULL: start at 0 as
 * fast.rs:217-218 does).  A proposal only starts searches it places within 20 of 0, and a
 * search from it that fails or ends beyond 20 runs again from 0, so every variant keeps the
 * outcome of the search from 0 (its root, or the unadjusted p where that search diverges);
 * the proposal changes the path, not the result.  The model must outlive its use by the
 * engine. */
int saige_engine_set_root_proposal(saige_engine* e, const saige_mlp* mlp);
/* Kernel of the per-variant sums for uint8 dosage rows (no counterpart in the reference: a
 * tuning and test control).  -1 (the default): chosen by timing the first batch of each
 * shape (at least 65,536 samples), else the int8 matrix-core form; 0: lanes over samples
 * (at most 3 covariates); 1: a lane per variant; 2: int8 matrix cores (exact integer sums of
 * the dosage bytes against a 70-bit fixed-point expansion of each feature).  Every form
 * computes the same sums; forms 0 and 1 add the double products in their own orders. */
int saige_engine_set_sums_form(saige_engine* e, int form);
/* Carrier-only SPA passes (no counterpart in the reference: a tuning and test control).  An
 * intercept-only engine holds a table of three sums over all samples as functions of one
 * variable (Chebyshev series fitted to the exact sums, accurate to their rounding); with it,
 * the K', K'' and K sums of a uint8 variant whose carriers (dosage > 0) are at most 35% of
 * the samples are taken over the carriers only, the samples of dosage 0 entering through
 * the table.  enable = 1 (the default) uses it, 0 sums over every sample; *has_table (may be
 * NULL) tells whether the engine has the table. */
int saige_engine_set_spa_split(saige_engine* e, int enable, int* has_table);
/* Copy src's tuned sums shapes and its sums-form and carrier-pass settings to dst (no
 * counterpart in the reference: for several engines over one null model, each on its own
 * stream and host thread, so that each batch is computed the same way, and so gives the same
 * bits, whichever engine takes it).  Both engines have the same n and covariate count. */
int saige_engine_copy_tuning(saige_engine* dst, saige_engine* src);

/* ------------------------------------------------------------------------------
 * SPA on one covariate-adjusted genotype vector.
 *   Replaces spa_binary_fast (fast.rs:207-242) / spa_binary (binary.rs:194-228).
 * ---------------------------------------------------------------------------- */
typedef struct {
  double pvalue;
  int32_t converged, is_spa, iters1, iters2;
  double root1, root2;
} saige_spa_result;

int saige_spa_binary(int device, const double* mu, const double* g, size_t n, double q,
                     double qinv, double pval_noadj, double tol, int fast,
                     saige_spa_result* out);

#ifdef __cplusplus
}
#endif
#endif /* SAIGE_GPU_H */
