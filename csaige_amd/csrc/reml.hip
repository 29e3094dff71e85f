// Fit-null on the GPU: AI-REML (ai_reml.rs:86-427), (X'WX)^-1 X'W (fit_null.rs:289-307),
// the variance ratio (variance_ratio.rs:87-246) and the whole fit_null::run flow
// (saige-cli/src/commands/fit_null.rs:199-394).
//
// All n-length work (Sigma solves, GRM products, projections, dot products) runs on the
// device; the host keeps only the p x p algebra (Cholesky of X' Sigma^-1 X, its solves)
// and the scalar tau update, exactly as the reference does them.
//
// Batching (iterate-identical to the reference): per AI-REML iteration the reference
// runs 1 + p + R separate PCG solves on the same Sigma (y~, the p columns of X, the R
// Hutchinson probes; ai_reml.rs:175-187, 237-258); here they are one block solve with
// k = 1 + p + R columns.  GRM * u of every probe does not change between iterations
// (ai_reml.rs:254 recomputes it) and is formed once.  The second-stage solves
// Sigma^-1 (GRM P y) (and, for quantitative traits, Sigma^-1 (P y / W)) form one block of
// 1 or 2 columns.
#include "common.hpp"
#include "rng.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

namespace saige {

int upload_repack(hipStream_t s, const uint8_t* bed_body, size_t bpm, size_t n_total,
                  const uint32_t* d_idx, size_t n, const uint32_t* marker_idx, size_t m,
                  uint8_t* d_dst, size_t row_bytes);
int upload_sample_idx(hipStream_t s, const uint32_t* sample_idx, size_t n, size_t n_total,
                      DevBuf& d);
int repack_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                       size_t n, const uint32_t* marker_idx, size_t m, uint8_t* d_dst,
                       size_t row_bytes);
int marker_stats_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                             size_t n, size_t m, double* ac, uint32_t* n_valid, double* af,
                             double* mac);
int grm_create_impl(int device, const uint8_t* bed_body, bool on_device, size_t bpm,
                    size_t n_total, const uint32_t* sample_idx, size_t n,
                    const uint32_t* marker_idx, size_t m, const double* af, int precision,
                    saige_grm** out);

// ---------------------------------------------------------------------------------
// p x p host algebra: CholeskyDecomp (saige-linalg/src/decomposition.rs:35-102).
// ---------------------------------------------------------------------------------
static bool chol(const std::vector<double>& a, size_t n, std::vector<double>& l) {
  l.assign(n * n, 0.0);
  for (size_t j = 0; j < n; ++j) {
    double sum = 0.0;
    for (size_t k = 0; k < j; ++k) sum += l[k * n + j] * l[k * n + j];
    double d = a[j * n + j] - sum;
    if (d <= 0.0) return false;
    l[j * n + j] = std::sqrt(d);
    for (size_t i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (size_t k = 0; k < j; ++k) s += l[k * n + i] * l[k * n + j];
      l[j * n + i] = (a[j * n + i] - s) / l[j * n + j];
    }
  }
  return true;
}
static void chol_solve(const std::vector<double>& l, size_t n, const double* b, double* x) {
  std::vector<double> y(n);
  for (size_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (size_t j = 0; j < i; ++j) s += l[j * n + i] * y[j];
    y[i] = (b[i] - s) / l[i * n + i];
  }
  for (size_t i = n; i-- > 0;) {
    double s = 0.0;
    for (size_t j = i + 1; j < n; ++j) s += l[i * n + j] * x[j];
    x[i] = (y[i] - s) / l[i * n + i];
  }
}

// ---------------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double inv_logit(double e) { return 1.0 / (1.0 + exp(-e)); }

// family.rs:55-90 working weights / residuals and the working response y~ = eta + r / W
__global__ void reml_prep_kernel(const double* __restrict__ y, const double* __restrict__ eta,
                                 const double* __restrict__ mu, int trait, size_t n,
                                 double* __restrict__ w, double* __restrict__ ytil) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = mu[i];
  double v = (trait == SAIGE_TRAIT_QUANTITATIVE) ? 1.0 : m * (1.0 - m);
  double wi = v > 1e-30 ? v : 1e-30;
  w[i] = wi;
  ytil[i] = eta[i] + (y[i] - m) / fmax(wi, 1e-30);
}

// eta = X alpha (dense.rs:102-113 order), mu = clamp(inv_logit(eta)) (family.rs:82-90)
__global__ void eta_mu_kernel(const double* __restrict__ x, size_t ldx, size_t p,
                              const double* __restrict__ alpha, int trait, size_t n,
                              double* __restrict__ eta, double* __restrict__ mu) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e = 0.0;
  for (size_t j = 0; j < p; ++j) e += x[i + j * ldx] * alpha[j];
  eta[i] = e;
  if (trait == SAIGE_TRAIT_QUANTITATIVE) {
    mu[i] = e;
  } else {
    double m = inv_logit(e);
    mu[i] = fmin(fmax(m, 1e-10), 1.0 - 1e-10);
  }
}

// out[:, c] = a[:, c] - sum_j six[:, j] * coef[j + p*c]  (ai_reml.rs:221-226, 246-251)
__global__ void project_kernel(const double* __restrict__ a, size_t lda,
                               const double* __restrict__ six, size_t lds, size_t p,
                               const double* __restrict__ coef, size_t n, size_t k,
                               double* __restrict__ out, size_t ldo) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  double v = a[i + c * lda];
  for (size_t j = 0; j < p; ++j) v -= six[i + j * lds] * coef[j + p * c];
  out[i + c * ldo] = v;
}

// out[:, c] = a[:, c] / max(w, 1e-30)
__global__ void div_w_kernel(const double* __restrict__ a, size_t lda,
                             const double* __restrict__ w, size_t n, size_t k,
                             double* __restrict__ out, size_t ldo) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  out[i + c * ldo] = a[i + c * lda] / fmax(w[i], 1e-30);
}

__global__ void fill_kernel(double* __restrict__ a, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

static inline dim3 grid1(size_t n) { return dim3((unsigned)ceil_div(n, 256)); }
static inline dim3 grid2(size_t n, size_t k) { return dim3((unsigned)ceil_div(n, 256), (unsigned)k); }

// IRLS working weights and response (ai_reml.rs:441-447) from the current (eta, mu):
// column j < p of out is w x_j and column p is w z.  With alpha null (eta, mu) are the
// start values from y (:436-439), otherwise eta = X alpha and mu its clamped inverse logit
// (:466-474).
__global__ void irls_prep_kernel(const double* __restrict__ x, size_t ldx, size_t p,
                                 const double* __restrict__ y, const double* __restrict__ alpha,
                                 int trait, size_t n, double* __restrict__ out, size_t ldo) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool quant = trait == SAIGE_TRAIT_QUANTITATIVE;
  double eta, mu;
  if (!alpha) {
    mu = quant ? y[i] : (y[i] + 0.5) / 2.0;
    eta = quant ? mu : log(mu / (1.0 - mu));
  } else {
    eta = 0.0;
    for (size_t j = 0; j < p; ++j) eta += x[i + j * ldx] * alpha[j];
    mu = quant ? eta : fmin(fmax(1.0 / (1.0 + exp(-eta)), 1e-10), 1.0 - 1e-10);
  }
  const double v = quant ? 1.0 : mu * (1.0 - mu);
  const double w = v > 1e-30 ? v : 1e-30;
  const double z = eta + (y[i] - mu) / fmax(w, 1e-30);
  for (size_t j = 0; j < p; ++j) out[i + j * ldo] = w * x[i + j * ldx];
  out[i + p * ldo] = w * z;
}

// IRLS for the GLM start (ai_reml.rs:431-482) on the device: per iteration one kernel for
// w and z, the deterministic Gram X'[WX | Wz], and the p x p Cholesky solve on the host.
static int fit_glm_irls_dev(hipStream_t s, const double* dX, size_t ldx, const double* dY,
                            size_t n, size_t p, int trait, int max_iter,
                            std::vector<double>& alpha, DevBuf& tmp, DevBuf& coef, DevBuf& part,
                            DevBuf& res) {
  const size_t ld = round_up(n, 64);
  SAIGE_TRY(tmp.ensure(ld * (p + 1) * 8));
  SAIGE_TRY(coef.ensure(std::max<size_t>(p, 8) * 8));
  std::vector<double> G(p * (p + 1)), xw(p * p), xz(p), an(p), l;
  alpha.assign(p, 0.0);
  for (int it = 0; it < max_iter; ++it) {
    irls_prep_kernel<<<dim3((unsigned)ceil_div(n, 256)), 256, 0, s>>>(
        dX, ldx, p, dY, it == 0 ? nullptr : coef.as<double>(), trait, n, tmp.as<double>(), ld);
    SAIGE_HIP(hipGetLastError());
    SAIGE_TRY(gram(s, dX, ldx, p, tmp.as<double>(), ld, p + 1, n, G.data(), part, res));
    for (size_t j = 0; j < p; ++j)
      for (size_t k = j; k < p; ++k) {
        xw[k * p + j] = G[j + p * k];  // sum_i x_j (w x_k)
        xw[j * p + k] = G[j + p * k];
      }
    for (size_t j = 0; j < p; ++j) xz[j] = G[j + p * p];
    if (!chol(xw, p, l)) {
      for (size_t i = 0; i < p; ++i) xw[i * p + i] += 1e-6;
      if (!chol(xw, p, l)) {
        set_error("GLM IRLS: X'WX not positive definite");
        return SAIGE_ERR_NOT_PD;
      }
    }
    chol_solve(l, p, xz.data(), an.data());
    double change = 0.0;
    for (size_t j = 0; j < p; ++j) change += std::fabs(an[j] - alpha[j]);
    alpha = an;
    SAIGE_HIP(hipMemcpyAsync(coef.p, alpha.data(), p * 8, hipMemcpyHostToDevice, s));
    if (change < 1e-8) break;
  }
  return SAIGE_OK;
}

struct ChaChaKey {
  uint32_t k[8];
};

// Rademacher probes from the ChaCha8 stream: thread b expands block b (stream words
// 16b..16b+15); word j is probe j / n, sample j % n, +1 when its sign bit is set.
__global__ void rademacher_kernel(ChaChaKey key, size_t n, size_t R, double* __restrict__ out,
                                  size_t ld) {
  const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t words = n * R;
  if (16 * b >= words) return;
  uint32_t w[16];
  chacha8_block(key.k, (uint64_t)b, w);
  for (int t = 0; t < 16; ++t) {
    const size_t j = 16 * b + t;
    if (j >= words) break;
    out[(j % n) + (j / n) * ld] = (w[t] >> 31) ? 1.0 : -1.0;
  }
}

struct RemlState {
  DevBuf x, y, eta, mu, w, rhs1, sol1, gu, py, gpy, pu, rhs2, sol2, tmp, coef, part, res, ones;
  DevBuf dres;  // scalars of one AI-REML step, read back together
  DevBuf glm, xinv, a_flat, at;  // IRLS work columns; (X'WX)^-1 and X'W-side products
  PcgWork pcg;
};

// fit_ai_reml with X (n x p, host col-major) and y (host); outputs stay on device in st.
// Per-chunk sums of the final log-likelihood terms (ai_reml.rs:393-414): binary
// y log(mu) + (1 - y) log(1 - mu) with mu clamped to [1e-15, 1 - 1e-15], quantitative
// (y - mu)^2; kDotChunk samples per block, a fixed reduction tree.
__global__ __launch_bounds__(kDotBlock) void loglik_terms_kernel(const double* __restrict__ y,
                                                                 const double* __restrict__ mu,
                                                                 size_t n, int trait,
                                                                 double* __restrict__ part) {
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, n);
  double acc = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    if (trait != SAIGE_TRAIT_QUANTITATIVE) {
      const double mi = fmin(fmax(mu[i], 1e-15), 1.0 - 1e-15);
      acc += y[i] * log(mi) + (1.0 - y[i]) * log(1.0 - mi);
    } else {
      const double d = y[i] - mu[i];
      acc += d * d;
    }
  }
  __shared__ double sh[kDotBlock];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static int fit_ai_reml_dev(saige_grm* h, ProductSink* sink, const double* y, const double* x,
                           size_t n, size_t p, int trait, const saige_ai_reml_config* cfg,
                           saige_ai_reml_result* res, RemlState& st) {
  hipStream_t s = h->stream;
  const size_t ld = round_up(n, 64);
  const size_t R = cfg->n_random_vectors;
  const bool fix_tau_e = (trait != SAIGE_TRAIT_QUANTITATIVE);
  const size_t k1 = 1 + p + R;
  *res = saige_ai_reml_result{};
  const uint64_t prod0 = sink->columns;

  std::vector<double> alpha;
  SAIGE_TRY(st.x.ensure(ld * p * 8));
  SAIGE_TRY(st.y.ensure(n * 8));
  SAIGE_TRY(st.eta.ensure(n * 8));
  SAIGE_TRY(st.mu.ensure(n * 8));
  SAIGE_TRY(st.w.ensure(n * 8));
  SAIGE_TRY(st.rhs1.ensure(ld * k1 * 8));
  SAIGE_TRY(st.sol1.ensure(ld * k1 * 8));
  SAIGE_TRY(st.gu.ensure(ld * std::max<size_t>(R, 1) * 8));
  SAIGE_TRY(st.py.ensure(ld * 8));
  SAIGE_TRY(st.gpy.ensure(ld * 8));
  SAIGE_TRY(st.pu.ensure(ld * std::max<size_t>(R, 1) * 8));
  SAIGE_TRY(st.rhs2.ensure(ld * 2 * 8));
  SAIGE_TRY(st.sol2.ensure(ld * 2 * 8));
  SAIGE_TRY(st.tmp.ensure(ld * (R + 2) * 8));
  SAIGE_TRY(st.coef.ensure(std::max<size_t>(p * std::max<size_t>(R, 2), 8) * 8));
  double* dX = st.x.as<double>();
  double* dY = st.y.as<double>();
  double* dEta = st.eta.as<double>();
  double* dMu = st.mu.as<double>();
  double* dW = st.w.as<double>();
  double* RHS1 = st.rhs1.as<double>();
  double* SOL1 = st.sol1.as<double>();
  double* GU = st.gu.as<double>();
  double* PY = st.py.as<double>();
  double* GPY = st.gpy.as<double>();
  double* PU = st.pu.as<double>();
  double* RHS2 = st.rhs2.as<double>();
  double* SOL2 = st.sol2.as<double>();
  double* TMP = st.tmp.as<double>();
  double* COEF = st.coef.as<double>();

  SAIGE_HIP(hipMemcpy2DAsync(dX, ld * 8, x, n * 8, n * 8, p, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(dY, y, n * 8, hipMemcpyHostToDevice, s));
  SAIGE_TRY(fit_glm_irls_dev(s, dX, ld, dY, n, p, trait, 25, alpha, st.glm, st.coef, st.part,
                             st.res));
  // X columns and the Rademacher probes are fixed right-hand sides (ai_reml.rs:121-130).
  SAIGE_HIP(hipMemcpy2DAsync(RHS1 + ld, ld * 8, x, n * 8, n * 8, p, hipMemcpyHostToDevice, s));
  if (R > 0) {
    // probe r, sample i is the sign bit of stream word r*n + i (ai_reml.rs:121-130),
    // generated on the device from the same key
    ChaCha8 rng(cfg->seed);
    ChaChaKey key;
    std::memcpy(key.k, rng.key(), sizeof(key.k));
    const size_t words = n * R;
    rademacher_kernel<<<(unsigned)ceil_div(ceil_div(words, 16), 256), 256, 0, s>>>(
        key, n, R, RHS1 + (1 + p) * ld, ld);
    SAIGE_HIP(hipGetLastError());
  }
  const double* U = RHS1 + (1 + p) * ld;
  // operand slices of the int8x2 form (the other forms ignore the request): three for every
  // product whose result enters the tau update directly or is computed once (GRM u, GRM P y,
  // the second-stage and variance-ratio solves); in the main block, the y~ and X columns take
  // two while the relative tau change is above kFineSwitch and three from then on, the probe
  // columns (whose solves only enter the Monte-Carlo trace estimate) two throughout.  Two
  // slices move tau by about 1e-5 relative; switching at 1e-3 keeps the two-slice iterations
  // well above that floor (at 1e-4 the fits took up to two extra iterations)
  constexpr int kFine = 3;
  constexpr double kFineSwitch = 1e-3;
  bool fine = h->precision != SAIGE_PREC_INT8X2;
  std::vector<uint8_t> blk_sl(k1, 0);  // per-column slice requests of the main block
  if (R > 0) {
    // constant GRM * u; the probes depend only on (seed, n, R), so fits of several traits
    // over the same samples share it
    sink->share_next(((uint64_t)cfg->seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)n << 20) ^ R);
    SAIGE_TRY(sink->product(U, ld, R, GU, ld, nullptr, nullptr, kFine));
  }

  // eta, mu from the GLM fit
  SAIGE_HIP(hipMemcpyAsync(COEF, alpha.data(), p * 8, hipMemcpyHostToDevice, s));
  eta_mu_kernel<<<grid1(n), 256, 0, s>>>(dX, ld, p, COEF, trait, n, dEta, dMu);
  SAIGE_HIP(hipGetLastError());

  double tau[2] = {1.0, 0.0};
  bool converged = false;
  uint32_t n_iter = 0;
  uint64_t pcg_iters = 0;
  uint32_t pcg_fail = 0;
  std::vector<saige_pcg_info> info(k1);
  std::vector<double> G((1 + p) * (p + R)), xsx(p * p), xsy(p), l, tcol(p), corr(p * std::max<size_t>(R, 2));
  std::vector<double> tr(std::max<size_t>(R, 1)), misc(4), sc1(R + 4);
  SAIGE_TRY(st.dres.ensure((R + 4) * 8));

  for (uint32_t iter = 0; iter < cfg->max_iter; ++iter) {
    n_iter = iter + 1;
    reml_prep_kernel<<<grid1(n), 256, 0, s>>>(dY, dEta, dMu, trait, n, dW, RHS1);
    SAIGE_HIP(hipGetLastError());
    const bool fine_iter = fine;
    // [y~ | X] on three slices once fine; the probe columns stay on the handle's form
    std::fill(blk_sl.begin(), blk_sl.begin() + 1 + p, (uint8_t)(fine_iter ? kFine : 0));
    SAIGE_TRY(pcg_sigma_device(h, sink, st.pcg, dW, tau, RHS1, ld, k1, cfg->pcg_tol,
                               cfg->pcg_max_iter, SOL1, ld, info.data(), 0, blk_sl.data()));
    for (size_t c = 0; c < k1; ++c) {
      pcg_iters += info[c].iterations;
      pcg_fail += info[c].converged ? 0 : 1;
    }
    const double* SIY = SOL1;
    const double* SIX = SOL1 + ld;
    const double* SIU = SOL1 + (1 + p) * ld;
    // [Sigma^-1 y~ | Sigma^-1 X]' [X | U]: X'Sigma^-1 y, X'Sigma^-1 X, (Sigma^-1 X)'u
    SAIGE_TRY(gram(s, SOL1, ld, 1 + p, RHS1 + ld, ld, p + R, n, G.data(), st.part, st.res));
    auto Gat = [&](size_t a, size_t b) { return G[a + (1 + p) * b]; };
    for (size_t j = 0; j < p; ++j)
      for (size_t k = j; k < p; ++k) {
        double d = Gat(1 + j, k);  // dot(x.col(k), sigma_inv_x[j]) (ai_reml.rs:194)
        xsx[k * p + j] = d;
        xsx[j * p + k] = d;
      }
    for (size_t j = 0; j < p; ++j) xsy[j] = Gat(0, j);
    if (!chol(xsx, p, l)) break;  // ai_reml.rs:207-213
    chol_solve(l, p, xsy.data(), alpha.data());
    for (size_t r = 0; r < R; ++r) {
      for (size_t j = 0; j < p; ++j) tcol[j] = Gat(1 + j, p + r);
      chol_solve(l, p, tcol.data(), corr.data() + p * r);
    }
    // alpha, eta, mu (ai_reml.rs:214-218)
    SAIGE_HIP(hipMemcpyAsync(COEF, alpha.data(), p * 8, hipMemcpyHostToDevice, s));
    eta_mu_kernel<<<grid1(n), 256, 0, s>>>(dX, ld, p, COEF, trait, n, dEta, dMu);
    // P y = Sigma^-1 y - Sigma^-1 X alpha (ai_reml.rs:221-226)
    project_kernel<<<grid2(n, 1), 256, 0, s>>>(SIY, ld, SIX, ld, p, COEF, n, 1, PY, ld);
    SAIGE_HIP(hipGetLastError());
    SAIGE_TRY(sink->product(PY, ld, 1, GPY, ld, nullptr, nullptr, kFine));
    // y'P GRM P y and the trace terms (GRM u)'P u in one read (dres[0], dres[1..R])
    double* DRES = st.dres.as<double>();
    SAIGE_TRY(col_dots_dev(s, PY, ld, GPY, ld, n, 1, DRES, st.part));
    if (R > 0) {
      // P u_r = Sigma^-1 u_r - Sigma^-1 X corr_r (ai_reml.rs:242-251)
      SAIGE_HIP(hipMemcpyAsync(COEF, corr.data(), p * R * 8, hipMemcpyHostToDevice, s));
      project_kernel<<<grid2(n, R), 256, 0, s>>>(SIU, ld, SIX, ld, p, COEF, n, R, PU, ld);
      SAIGE_HIP(hipGetLastError());
      SAIGE_TRY(col_dots_dev(s, GU, ld, PU, ld, n, R, DRES + 1, st.part));
    }
    SAIGE_TRY(read_doubles(s, DRES, 1 + R, sc1.data()));
    const double ypapy = sc1[0];
    double trace = 0.0;
    if (R > 0) {
      for (size_t r = 0; r < R; ++r) trace += sc1[1 + r];
      trace /= (double)R;
    }
    const double score_g = ypapy - trace;

    // second-stage block: Sigma^-1 (GRM P y) [and Sigma^-1 (P y / W)]
    size_t k2 = fix_tau_e ? 1 : 2;
    SAIGE_HIP(hipMemcpyAsync(RHS2, GPY, n * 8, hipMemcpyDeviceToDevice, s));
    if (!fix_tau_e) {
      div_w_kernel<<<grid2(n, 1), 256, 0, s>>>(PY, ld, dW, n, 1, RHS2 + ld, ld);
      SAIGE_HIP(hipGetLastError());
    }
    SAIGE_TRY(pcg_sigma_device(h, sink, st.pcg, dW, tau, RHS2, ld, k2, cfg->pcg_tol,
                               cfg->pcg_max_iter, SOL2, ld, info.data(), kFine));
    for (size_t c = 0; c < k2; ++c) {
      pcg_iters += info[c].iterations;
      pcg_fail += info[c].converged ? 0 : 1;
    }
    // corrections (Sigma^-1 X)' rhs2 -> chol.solve (ai_reml.rs:263-272, 306-315)
    std::vector<double> g2(p * 2);
    SAIGE_TRY(gram(s, SIX, ld, p, RHS2, ld, k2, n, g2.data(), st.part, st.res));
    for (size_t c = 0; c < k2; ++c) chol_solve(l, p, g2.data() + p * c, corr.data() + p * c);
    SAIGE_HIP(hipMemcpyAsync(COEF, corr.data(), p * k2 * 8, hipMemcpyHostToDevice, s));
    project_kernel<<<grid2(n, k2), 256, 0, s>>>(SOL2, ld, SIX, ld, p, COEF, n, k2, TMP, ld);
    SAIGE_HIP(hipGetLastError());
    // ai11 = dres[0]; without a fixed tau_e also ypa0py, ai00, ai01 = dres[1..3] and the
    // trace terms u'W^-1 P u = dres[4..4+R), all in one read
    SAIGE_TRY(col_dots_dev(s, GPY, ld, TMP, ld, n, 1, DRES, st.part));
    double score_e = 0.0, ai00 = 0.0, ai01 = 0.0;
    if (!fix_tau_e) {
      // TMP = [P GRM P y | P (P y / W) | U / W]; trace(P diag(1/W)) reuses P u
      // (ai_reml.rs:281-317 recomputes the same solves)
      SAIGE_TRY(col_dots_dev(s, PY, ld, RHS2 + ld, ld, n, 1, DRES + 1, st.part));
      SAIGE_TRY(col_dots_dev(s, RHS2 + ld, ld, TMP + ld, ld, n, 1, DRES + 2, st.part));
      SAIGE_TRY(col_dots_dev(s, RHS2 + ld, ld, TMP, ld, n, 1, DRES + 3, st.part));
      if (R > 0) {
        div_w_kernel<<<grid2(n, R), 256, 0, s>>>(U, ld, dW, n, R, TMP + 2 * ld, ld);
        SAIGE_HIP(hipGetLastError());
        SAIGE_TRY(col_dots_dev(s, TMP + 2 * ld, ld, PU, ld, n, R, DRES + 4, st.part));
      }
    }
    SAIGE_TRY(read_doubles(s, DRES, fix_tau_e ? 1 : 4 + R, sc1.data()));
    const double ai11 = sc1[0];
    if (!fix_tau_e) {
      double trw = 0.0;
      if (R > 0) {
        for (size_t r = 0; r < R; ++r) trw += sc1[4 + r];
        trw /= (double)R;
      }
      score_e = sc1[1] - trw;
      ai00 = sc1[2];
      ai01 = sc1[3];
    }
    // tau update with step halving (ai_reml.rs:329-369)
    double tn[2], max_change;
    if (fix_tau_e) {
      if (std::fabs(ai11) < 1e-30) break;
      double d1 = score_g / ai11;
      double step = 1.0;
      double tg = tau[1] + step * d1;
      while (tg < 0.0 && step > 1e-10) {
        step *= 0.5;
        tg = tau[1] + step * d1;
      }
      tg = std::max(tg, 1e-10);
      max_change = std::fabs(tg - tau[1]) / (std::fabs(tg) + std::fabs(tau[1]) + cfg->tol);
      tn[0] = 1.0;
      tn[1] = tg;
    } else {
      double det = ai00 * ai11 - ai01 * ai01;
      if (std::fabs(det) < 1e-30) break;
      double d0 = (ai11 * score_e - ai01 * score_g) / det;
      double d1 = (-ai01 * score_e + ai00 * score_g) / det;
      double step = 1.0;
      double t0 = tau[0] + step * d0, t1 = tau[1] + step * d1;
      while ((t0 < 0.0 || t1 < 0.0) && step > 1e-10) {
        step *= 0.5;
        t0 = tau[0] + step * d0;
        t1 = tau[1] + step * d1;
      }
      t0 = std::max(t0, 1e-10);
      t1 = std::max(t1, 1e-10);
      double c0 = std::fabs(t0 - tau[0]) / (std::fabs(t0) + std::fabs(tau[0]) + cfg->tol);
      double c1 = std::fabs(t1 - tau[1]) / (std::fabs(t1) + std::fabs(tau[1]) + cfg->tol);
      max_change = std::max(c0, c1);
      tn[0] = t0;
      tn[1] = t1;
    }
    tau[0] = tn[0];
    tau[1] = tn[1];
    if (max_change < kFineSwitch) fine = true;
    // a stop reached on two-slice block solves is confirmed by one more iteration on three
    if (max_change < cfg->tol && fine_iter) {
      converged = true;
      break;
    }
  }
  // final W and log-likelihood (ai_reml.rs:393-414): the per-sample terms summed on the
  // device in fixed-order chunks (the host adds the chunk sums in order), so a trait thread
  // of a joined fit does not hold the other traits' products back with an O(n) host loop
  double ll = 0.0;
  {
    const size_t nch = std::max<size_t>(1, ceil_div(n, kDotChunk));
    SAIGE_TRY(st.part.ensure(nch * 8));
    loglik_terms_kernel<<<(unsigned)nch, kDotBlock, 0, s>>>(dY, dMu, n, trait,
                                                             st.part.as<double>());
    SAIGE_HIP(hipGetLastError());
    std::vector<double> cs(nch);
    SAIGE_TRY(read_doubles(s, st.part.as<double>(), nch, cs.data()));
    double sum = 0.0;
    for (double c : cs) sum += c;
    if (trait != SAIGE_TRAIT_QUANTITATIVE)
      ll = sum;
    else
      ll = -0.5 * (double)n * std::log(2.0 * M_PI * tau[0]) - 0.5 * sum / tau[0];
  }
  // final working weights on device (family.rs:55-66)
  reml_prep_kernel<<<grid1(n), 256, 0, s>>>(dY, dEta, dMu, trait, n, dW, RHS1);
  SAIGE_HIP(hipGetLastError());
  res->tau[0] = tau[0];
  res->tau[1] = tau[1];
  res->iterations = n_iter;
  res->converged = converged ? 1 : 0;
  res->log_likelihood = ll;
  res->grm_products = sink->columns - prod0;
  res->pcg_iterations = pcg_iters;
  res->pcg_not_converged = pcg_fail;
  st.coef.ensure(std::max<size_t>(p, 8) * 8);
  SAIGE_HIP(hipMemcpyAsync(st.coef.p, alpha.data(), p * 8, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  return SAIGE_OK;
}

// (X'WX)^-1 X'W, flat buffer filled as fit_null.rs:293-305 does.
static int xvx_inv_xv_host(const double* x, size_t n, size_t p, const double* w, double* out) {
  std::vector<double> xvx(p * p), l, inv(p * p), e(p), col(p);
  for (size_t j = 0; j < p; ++j)
    for (size_t k = j; k < p; ++k) {
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) s += x[j * n + i] * w[i] * x[k * n + i];
      xvx[k * p + j] = s;
      xvx[j * p + k] = s;
    }
  if (!chol(xvx, p, l)) {
    set_error("X'WX is not positive definite");
    return SAIGE_ERR_NOT_PD;
  }
  for (size_t j = 0; j < p; ++j) {
    std::fill(e.begin(), e.end(), 0.0);
    e[j] = 1.0;
    chol_solve(l, p, e.data(), col.data());
    for (size_t i = 0; i < p; ++i) inv[j * p + i] = col[i];
  }
  for (size_t j = 0; j < p; ++j)
    for (size_t i = 0; i < n; ++i) {
      double v = 0.0;
      for (size_t k = 0; k < p; ++k) v += inv[k * p + j] * x[k * n + i] * w[i];
      out[j * n + i] = v;
    }
  return SAIGE_OK;
}

// (X'WX)^-1 X'W from the final W (fit_null.rs:289-307) on the device, in the flat layout
// the reference fills: a_flat[j*n + i] = sum_k inv[k*p + j] x_ik w_i.
__global__ void xvx_apply_kernel(const double* __restrict__ x, size_t ldx, size_t p,
                                 const double* __restrict__ w, const double* __restrict__ inv,
                                 size_t n, double* __restrict__ a_flat) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (size_t j = 0; j < p; ++j) {
    double v = 0.0;
    for (size_t k = 0; k < p; ++k) v += inv[k * p + j] * x[i + k * ldx] * w[i];
    a_flat[j * n + i] = v;
  }
}

// A^T as n x p col-major from the flat layout read as the col-major p x n view (the view
// estimate_variance_ratio takes, variance_ratio.rs:144-151): at[r*ld + i] = a_flat[i*p + r].
__global__ void at_gather_kernel(const double* __restrict__ a_flat, size_t n, size_t p,
                                 double* __restrict__ at, size_t ld) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t r = blockIdx.y;
  if (i >= n || r >= p) return;
  at[r * ld + i] = a_flat[i * p + r];
}

// columns j < p: w x_j
__global__ void wx_kernel(const double* __restrict__ x, size_t ldx, size_t p,
                          const double* __restrict__ w, size_t n, double* __restrict__ out,
                          size_t ldo) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i >= n || j >= p) return;
  out[i + j * ldo] = w[i] * x[i + j * ldx];
}

// st.a_flat and st.at from the device X and W; X'WX comes from the deterministic Gram, its
// inverse from the p x p Cholesky on the host.
static int xvx_inv_xv_dev(hipStream_t s, const double* dX, size_t ldx, const double* dW,
                          size_t n, size_t p, RemlState& st) {
  const size_t ld = round_up(n, 64);
  SAIGE_TRY(st.glm.ensure(ld * (p + 1) * 8));
  SAIGE_TRY(st.xinv.ensure(p * p * 8));
  SAIGE_TRY(st.a_flat.ensure(n * p * 8));
  SAIGE_TRY(st.at.ensure(ld * p * 8));
  wx_kernel<<<dim3((unsigned)ceil_div(n, 256), (unsigned)p), 256, 0, s>>>(dX, ldx, p, dW, n,
                                                                           st.glm.as<double>(), ld);
  SAIGE_HIP(hipGetLastError());
  std::vector<double> G(p * p), xvx(p * p), l, inv(p * p), e(p), col(p);
  SAIGE_TRY(gram(s, dX, ldx, p, st.glm.as<double>(), ld, p, n, G.data(), st.part, st.res));
  for (size_t j = 0; j < p; ++j)
    for (size_t k = j; k < p; ++k) {
      xvx[k * p + j] = G[j + p * k];
      xvx[j * p + k] = G[j + p * k];
    }
  if (!chol(xvx, p, l)) {
    set_error("X'WX is not positive definite");
    return SAIGE_ERR_NOT_PD;
  }
  for (size_t j = 0; j < p; ++j) {
    std::fill(e.begin(), e.end(), 0.0);
    e[j] = 1.0;
    chol_solve(l, p, e.data(), col.data());
    for (size_t i = 0; i < p; ++i) inv[j * p + i] = col[i];
  }
  SAIGE_HIP(hipMemcpyAsync(st.xinv.p, inv.data(), p * p * 8, hipMemcpyHostToDevice, s));
  xvx_apply_kernel<<<dim3((unsigned)ceil_div(n, 256)), 256, 0, s>>>(
      dX, ldx, p, dW, st.xinv.as<double>(), n, st.a_flat.as<double>());
  at_gather_kernel<<<dim3((unsigned)ceil_div(n, 256), (unsigned)p), 256, 0, s>>>(
      st.a_flat.as<double>(), n, p, st.at.as<double>(), ld);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

// ---------------------------------------------------------------------------------
// Variance ratio (variance_ratio.rs:87-246).
// ---------------------------------------------------------------------------------
// Flip to the minor allele per column where flip[c] (variance_ratio.rs:131-137).
__global__ void flip_kernel(double* __restrict__ g, size_t ld, size_t n, size_t k,
                            const int* __restrict__ flip) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k || !flip[c]) return;
  g[i + c * ld] = 2.0 - g[i + c * ld];
}
// g~ = g0 - X (A g0) (variance_ratio.rs:144-151; A g0 computed as p dots per column)
__global__ void adjust_kernel(const double* __restrict__ g0, const double* __restrict__ x,
                              size_t ld, size_t p, const double* __restrict__ ag, size_t n,
                              size_t k, double* __restrict__ gt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  double xa = 0.0;
  for (size_t j = 0; j < p; ++j) xa += x[i + j * ld] * ag[j + p * c];
  gt[i + c * ld] = g0[i + c * ld] - xa;
}
// weight * g~ / ac  (var2 summand without the last g~ factor)
__global__ void var2_weight_kernel(const double* __restrict__ gt, size_t ld,
                                   const double* __restrict__ mu, int trait,
                                   const double* __restrict__ ac, size_t n, size_t k,
                                   double* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  double m = mu[i], g = gt[i + c * ld];
  double wgt = trait == SAIGE_TRAIT_BINARY ? m * (1.0 - m) : (trait == SAIGE_TRAIT_QUANTITATIVE ? 1.0 : m);
  out[i + c * ld] = wgt * g / ac[c];
}
// decode marker rows (model order, padded rows) into f64 columns
__global__ void decode_rows_kernel(const uint8_t* __restrict__ rows, size_t row_bytes,
                                   const uint32_t* __restrict__ which, size_t n, size_t k,
                                   double* __restrict__ out, size_t ldo) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  uint32_t code = (rows[(size_t)which[c] * row_bytes + (i >> 2)] >> (2 * (i & 3))) & 3u;
  out[i + c * ldo] = code == 0u ? 2.0 : (code == 2u ? 1.0 : (code == 3u ? 0.0 : __longlong_as_double(0x7ff8000000000000ll)));
}

static double cal_cv(const std::vector<double>& v) {  // variance_ratio.rs:249-261
  size_t n = v.size();
  if (n <= 1) return INFINITY;
  double m = 0.0;
  for (double x : v) m += x;
  m /= (double)n;
  if (std::fabs(m) < 1e-30) return INFINITY;
  double var = 0.0;
  for (double x : v) var += (x - m) * (x - m);
  var /= ((double)n - 1.0);
  return (std::sqrt(var) / m) / (double)n;
}

// Column source: fills G (ld x k) with the raw dosages of candidate markers `ids`.
struct VrSource {
  const double* host_dosages = nullptr;  // n x m col-major (host)
  const uint8_t* d_rows = nullptr;       // device packed rows (model order)
  size_t row_bytes = 0;
};

static int vr_fill(hipStream_t s, const VrSource& src, size_t n, const std::vector<uint32_t>& ids,
                   double* G, size_t ld, DevBuf& d_ids) {
  size_t k = ids.size();
  if (src.host_dosages) {
    for (size_t c = 0; c < k; ++c)
      SAIGE_HIP(hipMemcpyAsync(G + c * ld, src.host_dosages + (size_t)ids[c] * n, n * 8,
                               hipMemcpyHostToDevice, s));
    return SAIGE_OK;
  }
  SAIGE_TRY(d_ids.ensure(k * 4));
  SAIGE_HIP(hipMemcpyAsync(d_ids.p, ids.data(), k * 4, hipMemcpyHostToDevice, s));
  decode_rows_kernel<<<grid2(n, k), 256, 0, s>>>(src.d_rows, src.row_bytes, d_ids.as<uint32_t>(),
                                                 n, k, G, ld);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

// compute_categorical_vr (variance_ratio.rs:263-287) with the default MAC bins of
// VarianceRatioConfig (:50-51): (upper bound, mean ratio) of every bin that holds a marker.
static void categorical_vr(const std::vector<double>& vr, const std::vector<double>& mac,
                           std::vector<std::pair<double, double>>& out) {
  static const double lo[8] = {0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 10.5, 20.5};
  static const double hi[8] = {1.5, 2.5, 3.5, 4.5, 5.5, 10.5, 20.5, INFINITY};
  out.clear();
  for (int b = 0; b < 8; ++b) {
    double sum = 0.0;
    size_t cnt = 0;
    for (size_t i = 0; i < vr.size(); ++i)
      if (mac[i] > lo[b] && mac[i] <= hi[b]) {
        sum += vr[i];
        ++cnt;
      }
    if (cnt) out.emplace_back(hi[b], sum / (double)cnt);
  }
}

static int estimate_vr_dev(saige_grm* h, ProductSink* sink, const VrSource& src,
                           const double* mac, size_t m,
                           const double* d_mu, const double* d_w, const double tau[2], int trait,
                           const double* d_x, size_t ldx, const double* d_at, size_t p,
                           const saige_vr_config* cfg, double* vr_out, uint32_t* n_used,
                           double* per_marker, RemlState& st,
                           std::vector<std::pair<double, double>>* cat = nullptr) {
  hipStream_t s = h->stream;
  const size_t n = h->g.n;
  const size_t ld = round_up(n, 64);
  ChaCha8 rng(cfg->seed);
  std::vector<uint32_t> order(m);
  std::iota(order.begin(), order.end(), 0u);
  rng.shuffle(order);
  // d_at: A^T as n x p col-major (leading dimension ld), element (r, i) of the col-major
  // p x n view of the flat A
  DevBuf G, GT, SG, WG, acb, agb, flipb, ids, onesb;
  SAIGE_TRY(onesb.ensure(ld * 8));
  fill_kernel<<<grid1(n), 256, 0, s>>>(onesb.as<double>(), n, 1.0);
  std::vector<double> vr, vr_mac;
  size_t target = cfg->n_markers, tested = 0, cursor = 0;
  bool done = false;
  while (!done && cursor < m) {
    size_t want = (tested < target ? target - tested : 0) + 2;
    want = std::max<size_t>(want, 4);
    std::vector<uint32_t> chunk;
    while (cursor < m && chunk.size() < want) {
      uint32_t id = order[cursor++];
      if (mac[id] < cfg->min_mac) continue;  // variance_ratio.rs:124-126
      chunk.push_back(id);
    }
    size_t k = chunk.size();
    if (k == 0) break;
    SAIGE_TRY(G.ensure(ld * k * 8));
    SAIGE_TRY(GT.ensure(ld * k * 8));
    SAIGE_TRY(SG.ensure(ld * k * 8));
    SAIGE_TRY(WG.ensure(ld * k * 8));
    SAIGE_TRY(acb.ensure(k * 8));
    SAIGE_TRY(agb.ensure(p * k * 8));
    SAIGE_TRY(flipb.ensure(k * 4));
    double* dG = G.as<double>();
    SAIGE_TRY(vr_fill(s, src, n, chunk, dG, ld, ids));
    std::vector<double> sums(k), acs(k), ag(p * k), v1(k), v2(k);
    SAIGE_TRY(col_dots(s, dG, ld, onesb.as<double>(), 0, n, k, sums.data(), st.part, st.res));
    std::vector<int> flip(k);
    for (size_t c = 0; c < k; ++c) flip[c] = (sums[c] / (2.0 * (double)n)) > 0.5;
    SAIGE_HIP(hipMemcpyAsync(flipb.p, flip.data(), k * 4, hipMemcpyHostToDevice, s));
    flip_kernel<<<grid2(n, k), 256, 0, s>>>(dG, ld, n, k, flipb.as<int>());
    SAIGE_TRY(col_dots(s, dG, ld, onesb.as<double>(), 0, n, k, acs.data(), st.part, st.res));
    SAIGE_TRY(gram(s, d_at, ld, p, dG, ld, k, n, ag.data(), st.part, st.res));
    SAIGE_HIP(hipMemcpyAsync(agb.p, ag.data(), p * k * 8, hipMemcpyHostToDevice, s));
    adjust_kernel<<<grid2(n, k), 256, 0, s>>>(dG, d_x, ldx, p, agb.as<double>(), n, k, GT.as<double>());
    SAIGE_HIP(hipGetLastError());
    std::vector<saige_pcg_info> info(k);
    SAIGE_TRY(pcg_sigma_device(h, sink, st.pcg, d_w, tau, GT.as<double>(), ld, k, 1e-5, 500,
                               SG.as<double>(), ld, info.data(), 3));
    SAIGE_TRY(col_dots(s, GT.as<double>(), ld, SG.as<double>(), ld, n, k, v1.data(), st.part, st.res));
    SAIGE_HIP(hipMemcpyAsync(acb.p, acs.data(), k * 8, hipMemcpyHostToDevice, s));
    var2_weight_kernel<<<grid2(n, k), 256, 0, s>>>(GT.as<double>(), ld, d_mu, trait, acb.as<double>(),
                                                   n, k, WG.as<double>());
    SAIGE_TRY(col_dots(s, WG.as<double>(), ld, GT.as<double>(), ld, n, k, v2.data(), st.part, st.res));
    // replay the sequential selection (variance_ratio.rs:120-212)
    for (size_t c = 0; c < k && !done; ++c) {
      double ac = acs[c];
      if (ac < 1.0) continue;
      double var1 = v1[c] / ac;
      double var2 = v2[c];
      if (var2 < 1e-30) continue;
      vr.push_back(var1 / var2);
      vr_mac.push_back(mac[chunk[c]]);
      ++tested;
      if (tested >= target) {
        if (cal_cv(vr) <= cfg->ratio_cv_cutoff) done = true;
        else target += 10;
      }
    }
  }
  if (vr.empty()) {
    set_error("No markers available for variance ratio estimation");
    return SAIGE_ERR_NO_VR_MARKERS;
  }
  double mean = 0.0;
  for (double v : vr) mean += v;
  mean /= (double)vr.size();
  *vr_out = mean;
  if (n_used) *n_used = (uint32_t)vr.size();
  if (per_marker) std::copy(vr.begin(), vr.end(), per_marker);
  if (cat) categorical_vr(vr, vr_mac, *cat);
  (void)d_x;
  return SAIGE_OK;
}

}  // namespace saige

using namespace saige;

extern "C" {

int saige_fit_ai_reml(saige_grm* h, const double* y, const double* x, size_t n, size_t p,
                      int trait_type, const saige_ai_reml_config* cfg,
                      saige_ai_reml_result* result, double* mu, double* w, double* residuals,
                      double* eta, double* alpha) {
  clear_error();
  SAIGE_ARG(h && y && x && cfg && result, "null pointer");
  SAIGE_ARG(n == h->g.n, "y length != GRM samples");
  SAIGE_ARG(p >= 1, "p >= 1");
  SAIGE_ARG(trait_type >= 0 && trait_type <= 2, "trait_type");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  RemlState st;
  DirectSink sink(h);
  SAIGE_TRY(fit_ai_reml_dev(h, &sink, y, x, n, p, trait_type, cfg, result, st));
  hipStream_t s = h->stream;
  std::vector<double> hm(n), hw(n);
  SAIGE_HIP(hipMemcpyAsync(hm.data(), st.mu.p, n * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipMemcpyAsync(hw.data(), st.w.p, n * 8, hipMemcpyDeviceToHost, s));
  if (eta) SAIGE_HIP(hipMemcpyAsync(eta, st.eta.p, n * 8, hipMemcpyDeviceToHost, s));
  if (alpha) SAIGE_HIP(hipMemcpyAsync(alpha, st.coef.p, p * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  if (mu) std::copy(hm.begin(), hm.end(), mu);
  if (w) std::copy(hw.begin(), hw.end(), w);
  if (residuals)
    for (size_t i = 0; i < n; ++i) residuals[i] = y[i] - hm[i];
  return SAIGE_OK;
}

int saige_xvx_inv_xv(const double* x, size_t n, size_t p, const double* w, double* out) {
  clear_error();
  SAIGE_ARG(x && w && out && p >= 1, "null pointer");
  return xvx_inv_xv_host(x, n, p, w, out);
}

int saige_estimate_variance_ratio(saige_grm* h, const double* dosages, const double* mac, size_t m,
                                  const double* mu, const double* w, const double tau[2],
                                  int trait_type, const double* x, const double* xvx_inv_xv,
                                  size_t p, const saige_vr_config* cfg, double* variance_ratio,
                                  uint32_t* n_used, double* per_marker_vr) {
  clear_error();
  SAIGE_ARG(h && mac && mu && w && tau && x && xvx_inv_xv && cfg && variance_ratio, "null pointer");
  SAIGE_ARG(m == 0 || dosages, "dosages");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  const size_t n = h->g.n, ld = round_up(n, 64);
  hipStream_t s = h->stream;
  DevBuf dmu, dw, dx;
  SAIGE_TRY(dmu.ensure(n * 8));
  SAIGE_TRY(dw.ensure(n * 8));
  SAIGE_TRY(dx.ensure(ld * p * 8));
  SAIGE_HIP(hipMemcpyAsync(dmu.p, mu, n * 8, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(dw.p, w, n * 8, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpy2DAsync(dx.p, ld * 8, x, n * 8, n * 8, p, hipMemcpyHostToDevice, s));
  VrSource src;
  src.host_dosages = dosages;
  RemlState st;
  if (m == 0) {
    set_error("No markers available for variance ratio estimation");
    return SAIGE_ERR_NO_VR_MARKERS;
  }
  // A^T as n x p col-major: element (r, i) of the col-major p x n view is xvx_inv_xv[i*p + r]
  DevBuf d_at;
  SAIGE_TRY(d_at.ensure(ld * p * 8));
  std::vector<double> at(n * p);
  for (size_t i = 0; i < n; ++i)
    for (size_t r = 0; r < p; ++r) at[r * n + i] = xvx_inv_xv[i * p + r];
  SAIGE_HIP(hipMemcpy2DAsync(d_at.p, ld * 8, at.data(), n * 8, n * 8, p, hipMemcpyHostToDevice, s));
  DirectSink sink(h);
  return estimate_vr_dev(h, &sink, src, mac, m, dmu.as<double>(), dw.as<double>(), tau, trait_type,
                         dx.as<double>(), ld, d_at.as<double>(), p, cfg, variance_ratio, n_used,
                         per_marker_vr, st);
}

}  // extern "C"

namespace saige {

// QC and the VR/GRM marker split shared by every trait of a fit (fit_null.rs:199-264).
struct MarkerSplit {
  std::vector<uint32_t> grm_ids, vr_ids;
  std::vector<double> grm_af, vr_mac;
};

static int qc_split(int device, const uint8_t* bed_body, bool on_device, size_t bpm, size_t n_total,
                    size_t m_total, const uint32_t* sample_idx, size_t n,
                    const saige_fit_null_config* cfg, MarkerSplit& ms, uint8_t* marker_class) {
  // QC over all markers (fit_null.rs:209-222)
  std::vector<double> ac(m_total), af(m_total), mac(m_total);
  std::vector<uint32_t> nv(m_total);
  if (on_device) {
    SAIGE_ARG(bpm >= (n_total + 3) / 4, "bytes_per_marker < ceil(n_total/4)");
    if (!sample_idx) SAIGE_ARG(n == n_total, "n must equal n_total without a subset");
    SAIGE_TRY(use_device(device));
    hipStream_t qs;
    SAIGE_HIP(hipStreamCreate(&qs));
    DevBuf d_idx;
    int qrc = upload_sample_idx(qs, sample_idx, n, n_total, d_idx);
    if (qrc == SAIGE_OK)
      qrc = marker_stats_from_device(qs, bed_body, bpm, d_idx.as<uint32_t>(), n, m_total, ac.data(),
                                     nv.data(), af.data(), mac.data());
    (void)hipStreamDestroy(qs);
    SAIGE_TRY(qrc);
  } else {
    SAIGE_TRY(saige_plink_marker_stats(device, bed_body, bpm, n_total, m_total, sample_idx, n,
                                       ac.data(), nv.data(), af.data(), mac.data()));
  }
  std::vector<uint32_t> passing;
  for (size_t j = 0; j < m_total; ++j) {
    double miss = 1.0 - ((double)nv[j] / (double)n);
    if (af[j] >= cfg->min_maf && af[j] <= 1.0 - cfg->min_maf && miss <= cfg->max_missing_rate)
      passing.push_back((uint32_t)j);
  }
  if (passing.empty()) {
    set_error("no marker passes QC");
    return SAIGE_ERR_NO_MARKERS;
  }
  // VR candidates: 1000 draws of gen_range(0..n_passing), sort, dedup, MAC >= 20 (:230-243)
  const size_t np = passing.size();
  std::vector<uint8_t> in_vr(np, 0);
  {
    ChaCha8 rng(cfg->reml.seed);
    std::vector<uint64_t> cand(1000);
    for (auto& c : cand) c = rng.gen_range64(np);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    for (uint64_t c : cand)
      if (mac[passing[c]] >= 20.0) in_vr[c] = 1;
  }
  if (marker_class) std::fill(marker_class, marker_class + m_total, 0);
  for (size_t i = 0; i < np; ++i) {
    uint32_t j = passing[i];
    if (in_vr[i]) {
      ms.vr_ids.push_back(j);
      ms.vr_mac.push_back(mac[j]);
      if (marker_class) marker_class[j] = 2;
    } else {
      ms.grm_ids.push_back(j);
      ms.grm_af.push_back(af[j]);
      if (marker_class) marker_class[j] = 1;
    }
  }
  return SAIGE_OK;
}

// One trait after the GRM exists: AI-REML, XVX^-1X'V and the variance ratio
// (fit_null.rs:267-373), every GRM product through `sink`.  vr_rows: the VR markers as
// packed rows in model order (shared, read-only).
static int fit_trait(saige_grm* h, ProductSink* sink, const double* y, const double* x, size_t p,
                     int trait_type, const saige_fit_null_config* cfg, const MarkerSplit& ms,
                     const uint8_t* vr_rows, saige_fit_null_result* result,
                     double* per_marker_vr, RemlState& st) {
  using clk = std::chrono::steady_clock;
  const size_t n = h->g.n;
  hipStream_t s = h->stream;
  const size_t ld = round_up(n, 64);
  auto t1 = clk::now();
  SAIGE_TRY(fit_ai_reml_dev(h, sink, y, x, n, p, trait_type, &cfg->reml, &result->reml, st));
  auto t2 = clk::now();
  SAIGE_TRY(xvx_inv_xv_dev(s, st.x.as<double>(), ld, st.w.as<double>(), n, p, st));
  double vr = 1.0;
  uint32_t used = 0;
  if (!ms.vr_ids.empty()) {
    VrSource src;
    src.d_rows = vr_rows;
    src.row_bytes = h->g.row_bytes;
    saige_vr_config vc{cfg->n_markers_vr, 20.0, 0.001, cfg->reml.seed};
    std::vector<std::pair<double, double>> cat;
    SAIGE_TRY(estimate_vr_dev(h, sink, src, ms.vr_mac.data(), ms.vr_ids.size(),
                              st.mu.as<double>(), st.w.as<double>(), result->reml.tau, trait_type,
                              st.x.as<double>(), ld, st.at.as<double>(), p, &vc, &vr, &used,
                              per_marker_vr, st, cfg->use_categorical_vr ? &cat : nullptr));
    result->n_categorical_vr = (uint32_t)cat.size();
    for (size_t b = 0; b < cat.size() && b < 8; ++b) {
      result->categorical_vr_upper[b] = cat[b].first;
      result->categorical_vr[b] = cat[b].second;
    }
  }
  auto t3 = clk::now();
  result->variance_ratio = vr;
  result->vr_markers_used = used;
  result->n_grm_markers = (uint32_t)ms.grm_ids.size();
  result->n_vr_markers = (uint32_t)ms.vr_ids.size();
  result->seconds_reml = std::chrono::duration<double>(t2 - t1).count();
  result->seconds_vr = std::chrono::duration<double>(t3 - t2).count();
  result->reml.grm_products = sink->columns;
  return SAIGE_OK;
}

// mu, XVX^-1X'V (p x n, a_flat) and alpha of a fitted trait to the caller's host arrays.  A
// joined fit calls this after the trait has left the batcher: the copies into pageable memory
// hold the calling thread, which the other traits would otherwise wait for.
static int copy_trait_outputs(saige_grm* h, const RemlState& st, size_t p, double* mu_out,
                              double* xvx_inv_xv, double* alpha_out) {
  const size_t n = h->g.n;
  hipStream_t s = h->stream;
  if (mu_out) SAIGE_HIP(hipMemcpyAsync(mu_out, st.mu.p, n * 8, hipMemcpyDeviceToHost, s));
  if (xvx_inv_xv)
    SAIGE_HIP(hipMemcpyAsync(xvx_inv_xv, st.a_flat.p, n * p * 8, hipMemcpyDeviceToHost, s));
  if (alpha_out) SAIGE_HIP(hipMemcpyAsync(alpha_out, st.coef.p, p * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  return SAIGE_OK;
}

// VR markers resident as packed rows in model order (decoded per chunk by the VR solves).
static int vr_rows_upload(saige_grm* h, const uint8_t* bed_body, bool on_device, size_t bpm,
                          size_t n_total, const uint32_t* sample_idx, size_t n,
                          const MarkerSplit& ms, DevBuf& rows) {
  if (ms.vr_ids.empty()) return SAIGE_OK;
  hipStream_t s = h->stream;
  DevBuf d_idx;
  const size_t row_bytes = h->g.row_bytes;
  SAIGE_TRY(rows.ensure(ms.vr_ids.size() * row_bytes));
  SAIGE_TRY(upload_sample_idx(s, sample_idx, n, n_total, d_idx));
  if (on_device)
    SAIGE_TRY(repack_from_device(s, bed_body, bpm, d_idx.as<uint32_t>(), n, ms.vr_ids.data(),
                                 ms.vr_ids.size(), rows.as<uint8_t>(), row_bytes));
  else
    SAIGE_TRY(upload_repack(s, bed_body, bpm, n_total, d_idx.as<uint32_t>(), n, ms.vr_ids.data(),
                            ms.vr_ids.size(), rows.as<uint8_t>(), row_bytes));
  SAIGE_HIP(hipStreamSynchronize(s));
  return SAIGE_OK;
}

static int fit_null_impl(int device, const uint8_t* bed_body, bool on_device, size_t bpm,
                         size_t n_total, size_t m_total, const uint32_t* sample_idx, size_t n,
                         const double* y, const double* x, size_t p, int trait_type,
                         const saige_fit_null_config* cfg, saige_fit_null_result* result,
                         double* mu_out, double* xvx_inv_xv, uint8_t* marker_class,
                         double* alpha_out, double* per_marker_vr) {
  SAIGE_ARG(bed_body && y && x && cfg && result, "null pointer");
  SAIGE_ARG(p >= 1 && n > 0, "dimensions");
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  *result = saige_fit_null_result{};
  MarkerSplit ms;
  SAIGE_TRY(qc_split(device, bed_body, on_device, bpm, n_total, m_total, sample_idx, n, cfg, ms,
                     marker_class));
  auto t1 = clk::now();
  saige_grm* h = nullptr;
  SAIGE_TRY(grm_create_impl(device, bed_body, on_device, bpm, n_total, sample_idx, n,
                            ms.grm_ids.data(), ms.grm_ids.size(), ms.grm_af.data(), cfg->precision,
                            &h));
  int rc;
  {
    std::lock_guard<std::recursive_mutex> lk(h->mu);
    DevBuf rows;
    rc = vr_rows_upload(h, bed_body, on_device, bpm, n_total, sample_idx, n, ms, rows);
    DirectSink sink(h);
    RemlState st;
    if (rc == SAIGE_OK)
      rc = fit_trait(h, &sink, y, x, p, trait_type, cfg, ms, rows.as<uint8_t>(), result,
                     per_marker_vr, st);
    if (rc == SAIGE_OK) rc = copy_trait_outputs(h, st, p, mu_out, xvx_inv_xv, alpha_out);
  }
  saige_grm_destroy(h);
  if (rc == SAIGE_OK) {
    // the REML clock of fit_trait starts after the GRM build; charge the build to it
    result->seconds_qc = std::chrono::duration<double>(t1 - t0).count();
    result->seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
    result->seconds_reml = result->seconds_total - result->seconds_qc - result->seconds_vr;
  }
  return rc;
}

// ---------------------------------------------------------------------------------
// Several phenotypes over one GRM (the sample set and covariates are shared).
// ---------------------------------------------------------------------------------
// Every trait runs its own fit (its own W, tau, PCG iterations and stop rules) on its own
// host thread; the GRM products they request meet in a ProductBatcher, which waits until
// every trait still running has a product pending, runs one multi-column product over all
// of their columns and applies each request's own Sigma epilogue.  The per-column
// arithmetic of the product does not depend on which other columns share it, so each trait
// follows the iteration of a single-trait fit.

// out = tau0 * v / max(w, 1e-30) + tau1 * g  (or out = g without w)
__global__ void sigma_epilogue_kernel(const double* __restrict__ g, size_t ldg,
                                      const double* __restrict__ v, size_t ldv,
                                      const double* __restrict__ w, double tau0, double tau1,
                                      size_t n, size_t k, double* __restrict__ out, size_t ldo) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  const double y = g[i + c * ldg];
  out[i + c * ldo] = w ? tau0 * v[i + c * ldv] / fmax(w[i], 1e-30) + tau1 * y : y;
}

class ProductBatcher {
 public:
  ProductBatcher(saige_grm* h, size_t clients) : h_(h), live_(clients) {}

  struct Request {
    const double* v;
    size_t ldv, k;
    double* out;
    size_t ldo;
    const double* w;
    double tau[2];
    int loco = -1;  // chromosome left out of this request's GRM (-1: none)
    int slices = 0; // int8 operand slices asked for (0: the handle's own form)
    const uint8_t* col_slices = nullptr;  // or per column (k entries)
    uint64_t share = 0;  // nonzero: same input as every other request with this key
    int rc = SAIGE_OK;
    bool done = false;
  };

  // Blocks until the request's product has been enqueued on the handle's stream.
  int submit(Request& r) {
    std::unique_lock<std::mutex> lk(mu_);
    r.done = false;
    pending_.push_back(&r);
    if (pending_.size() == live_) run_locked();
    else cv_.wait(lk, [&] { return r.done; });
    return r.rc;
  }
  // A client has no more requests.
  void leave() {
    std::unique_lock<std::mutex> lk(mu_);
    --live_;
    if (live_ > 0 && pending_.size() == live_) run_locked();
  }
  uint64_t joined_products = 0;  // multi-column products run
  uint64_t joined_columns = 0;
  uint64_t shared_columns = 0;  // columns copied from an identical request instead

 private:
  void run_locked() {
    const int rc = run(pending_);
    for (Request* q : pending_) {
      q->rc = rc;
      q->done = true;
    }
    pending_.clear();
    cv_.notify_all();
  }
  int run(const std::vector<Request*>& all) {
    ++joined_products;
    // requests with the same share key, slices and left-out chromosome are computed once
    std::vector<Request*> reqs;
    std::vector<std::pair<Request*, Request*>> dup;  // (duplicate, the request computing it)
    for (Request* q : all) {
      Request* first = nullptr;
      if (q->share)
        for (Request* u : reqs)
          if (u->share == q->share && u->k == q->k && u->loco == q->loco &&
              u->slices == q->slices && !u->col_slices && !q->col_slices && !u->w && !q->w) {
            first = u;
            break;
          }
      if (first) dup.emplace_back(q, first);
      else reqs.push_back(q);
    }
    // per-column left-out chromosomes and operand slices of the joined product
    std::vector<int> codes;
    std::vector<uint8_t> sl;
    bool loco = false, slices = false;
    for (const Request* q : reqs) {
      codes.insert(codes.end(), q->k, q->loco);
      if (q->col_slices) {
        sl.insert(sl.end(), q->col_slices, q->col_slices + q->k);
        slices = true;
      } else {
        sl.insert(sl.end(), q->k, (uint8_t)q->slices);
        slices = slices || q->slices != 0;
      }
      loco = loco || q->loco >= 0;
    }
    h_->loco_req = loco ? codes.data() : nullptr;
    h_->slice_req = slices ? sl.data() : nullptr;
    const int rc = run_joined(reqs);
    h_->loco_req = nullptr;
    h_->slice_req = nullptr;
    if (rc != SAIGE_OK) return rc;
    for (const auto& d : dup) {  // plain products (no Sigma form): copy the shared result
      const size_t n = h_->g.n;
      SAIGE_HIP(hipMemcpy2DAsync(d.first->out, d.first->ldo * 8, d.second->out,
                                 d.second->ldo * 8, n * 8, d.first->k, hipMemcpyDeviceToDevice,
                                 h_->stream));
      shared_columns += d.first->k;
    }
    return SAIGE_OK;
  }
  int run_joined(const std::vector<Request*>& reqs) {
    if (reqs.size() == 1) {
      const Request& q = *reqs[0];
      joined_columns += q.k;
      return grm_product(h_, q.v, q.ldv, q.k, q.out, q.ldo, q.w, q.w ? q.tau : nullptr);
    }
    const size_t n = h_->g.n;
    hipStream_t s = h_->stream;
    size_t K = 0;
    for (const Request* q : reqs) K += q->k;
    joined_columns += K;
    SAIGE_TRY(vin_.ensure(n * K * 8));
    SAIGE_TRY(vout_.ensure(n * K * 8));
    double* vin = vin_.as<double>();
    double* vout = vout_.as<double>();
    size_t off = 0;
    for (const Request* q : reqs) {
      SAIGE_HIP(hipMemcpy2DAsync(vin + off * n, n * 8, q->v, q->ldv * 8, n * 8, q->k,
                                 hipMemcpyDeviceToDevice, s));
      off += q->k;
    }
    SAIGE_TRY(grm_product(h_, vin, n, K, vout, n, nullptr, nullptr));
    off = 0;
    for (const Request* q : reqs) {
      sigma_epilogue_kernel<<<dim3((unsigned)ceil_div(n, 256), (unsigned)q->k), 256, 0, s>>>(
          vout + off * n, n, q->v, q->ldv, q->w, q->tau[0], q->tau[1], n, q->k, q->out, q->ldo);
      off += q->k;
    }
    SAIGE_HIP(hipGetLastError());
    return SAIGE_OK;
  }

  saige_grm* h_;
  size_t live_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<Request*> pending_;
  DevBuf vin_, vout_;
};

struct BatchedSink : ProductSink {
  ProductBatcher* b;
  int loco = -1;
  uint64_t share = 0;  // key of the next product (share_next), then cleared
  explicit BatchedSink(ProductBatcher* bb) : b(bb) {}
  void share_next(uint64_t key) override { share = key; }
  int product(const double* v, size_t ldv, size_t k, double* out, size_t ldo, const double* w,
              const double* tau, int slices = 0, const uint8_t* col_slices = nullptr) override {
    const uint64_t key = share;
    share = 0;
    if (k == 0) return SAIGE_OK;
    columns += k;
    ProductBatcher::Request r;
    r.share = key;
    r.slices = slices;
    r.col_slices = col_slices;
    r.v = v;
    r.ldv = ldv;
    r.k = k;
    r.out = out;
    r.ldo = ldo;
    r.w = w;
    r.tau[0] = tau ? tau[0] : 0.0;
    r.tau[1] = tau ? tau[1] : 0.0;
    r.loco = loco;
    return b->submit(r);
  }
};

static int fit_null_multi_impl(int device, const uint8_t* bed_body, bool on_device, size_t bpm,
                               size_t n_total, size_t m_total, const uint32_t* sample_idx,
                               size_t n, const double* y, size_t n_traits, const double* x,
                               size_t p, const int* trait_types, const saige_fit_null_config* cfg,
                               saige_fit_null_result* results, double* mu_out,
                               double* xvx_inv_xv, uint8_t* marker_class, double* alpha_out,
                               double* per_marker_vr) {
  SAIGE_ARG(bed_body && y && x && cfg && results && trait_types, "null pointer");
  SAIGE_ARG(p >= 1 && n > 0 && n_traits >= 1 && n_traits <= 64, "dimensions");
  for (size_t t = 0; t < n_traits; ++t)
    SAIGE_ARG(trait_types[t] >= 0 && trait_types[t] <= 2, "trait_type");
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  for (size_t t = 0; t < n_traits; ++t) results[t] = saige_fit_null_result{};
  MarkerSplit ms;
  SAIGE_TRY(qc_split(device, bed_body, on_device, bpm, n_total, m_total, sample_idx, n, cfg, ms,
                     marker_class));
  auto t1 = clk::now();
  saige_grm* h = nullptr;
  SAIGE_TRY(grm_create_impl(device, bed_body, on_device, bpm, n_total, sample_idx, n,
                            ms.grm_ids.data(), ms.grm_ids.size(), ms.grm_af.data(), cfg->precision,
                            &h));
  int rc;
  {
    std::lock_guard<std::recursive_mutex> lk(h->mu);
    DevBuf rows;
    rc = vr_rows_upload(h, bed_body, on_device, bpm, n_total, sample_idx, n, ms, rows);
    if (rc == SAIGE_OK) {
      ProductBatcher batcher(h, n_traits);
      std::vector<int> rcs(n_traits, SAIGE_OK);
      std::vector<std::string> errs(n_traits);
      std::vector<std::thread> threads;
      for (size_t t = 0; t < n_traits; ++t) {
        threads.emplace_back([&, t]() {
          BatchedSink sink(&batcher);
          RemlState st;
          int r = hipSetDevice(h->device) == hipSuccess ? SAIGE_OK : SAIGE_ERR_DEVICE;
          if (r == SAIGE_OK)
            r = fit_trait(h, &sink, y + t * n, x, p, trait_types[t], cfg, ms, rows.as<uint8_t>(),
                          &results[t], per_marker_vr ? per_marker_vr + t * m_total : nullptr, st);
          batcher.leave();
          if (r == SAIGE_OK)
            r = copy_trait_outputs(h, st, p, mu_out ? mu_out + t * n : nullptr,
                                   xvx_inv_xv ? xvx_inv_xv + t * p * n : nullptr,
                                   alpha_out ? alpha_out + t * p : nullptr);
          if (r != SAIGE_OK) errs[t] = saige_last_error();
          rcs[t] = r;
        });
      }
      for (auto& th : threads) th.join();
      for (size_t t = 0; t < n_traits && rc == SAIGE_OK; ++t)
        if (rcs[t] != SAIGE_OK) {
          rc = rcs[t];
          set_error("trait %zu: %s", t, errs[t].c_str());
        }
    }
  }
  saige_grm_destroy(h);
  if (rc == SAIGE_OK) {
    const double total = std::chrono::duration<double>(clk::now() - t0).count();
    for (size_t t = 0; t < n_traits; ++t) {
      results[t].seconds_qc = std::chrono::duration<double>(t1 - t0).count();
      results[t].seconds_total = total;
    }
  }
  return rc;
}

// Leave-one-chromosome-out refits (glmm/loco.rs:54-114): one AI-REML per chromosome code
// 0..n_chrom-1, each on the GRM without that chromosome's markers.  The refits run on host
// threads and their products meet in a ProductBatcher: one joined product streams the
// genotypes once for every refit still iterating, each column masked for its chromosome.
static int fit_loco_impl(saige_grm* h, const double* y, const double* x, size_t n, size_t p,
                         int trait_type, const saige_ai_reml_config* cfg, size_t n_chrom,
                         saige_ai_reml_result* results, double* mu, double* w,
                         double* residuals, double* alpha) {
  ProductBatcher batcher(h, n_chrom);
  std::vector<int> rcs(n_chrom, SAIGE_OK);
  std::vector<std::string> errs(n_chrom);
  std::vector<std::thread> threads;
  for (size_t c = 0; c < n_chrom; ++c) {
    threads.emplace_back([&, c]() {
      BatchedSink sink(&batcher);
      sink.loco = (int)c;
      int r = hipSetDevice(h->device) == hipSuccess ? SAIGE_OK : SAIGE_ERR_DEVICE;
      RemlState st;
      if (r == SAIGE_OK)
        r = fit_ai_reml_dev(h, &sink, y, x, n, p, trait_type, cfg, &results[c], st);
      batcher.leave();
      if (r == SAIGE_OK) {
        hipStream_t s = h->stream;
        std::vector<double> hm(n);
        hipError_t e = hipMemcpyAsync(hm.data(), st.mu.p, n * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && w) e = hipMemcpyAsync(w + c * n, st.w.p, n * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && alpha)
          e = hipMemcpyAsync(alpha + c * p, st.coef.p, p * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
          set_error("LOCO outputs: %s", hipGetErrorString(e));
          r = SAIGE_ERR_DEVICE;
        } else {
          if (mu) std::copy(hm.begin(), hm.end(), mu + c * n);
          if (residuals)
            for (size_t i = 0; i < n; ++i) residuals[c * n + i] = y[i] - hm[i];
        }
      }
      if (r != SAIGE_OK) errs[c] = saige_last_error();
      rcs[c] = r;
    });
  }
  for (auto& th : threads) th.join();
  for (size_t c = 0; c < n_chrom; ++c)
    if (rcs[c] != SAIGE_OK) {
      set_error("chromosome code %zu: %s", c, errs[c].c_str());
      return rcs[c];
    }
  return SAIGE_OK;
}

}  // namespace saige

extern "C" {

int saige_fit_ai_reml_loco(saige_grm* h, const double* y, const double* x, size_t n, size_t p,
                           int trait_type, const saige_ai_reml_config* cfg, size_t n_chrom,
                           saige_ai_reml_result* results, double* mu, double* w,
                           double* residuals, double* alpha) {
  clear_error();
  SAIGE_ARG(h && y && x && cfg && results, "null pointer");
  SAIGE_ARG(n == h->g.n, "y length != GRM samples");
  SAIGE_ARG(p >= 1, "p >= 1");
  SAIGE_ARG(trait_type >= 0 && trait_type <= 2, "trait_type");
  SAIGE_ARG(n_chrom >= 1 && n_chrom <= 64, "1 <= n_chrom <= 64");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_ARG(h->loco_n_chrom > 0 && (size_t)h->loco_n_chrom == n_chrom,
            "n_chrom must equal the number of chromosome codes set on the handle");
  SAIGE_HIP(hipSetDevice(h->device));
  return fit_loco_impl(h, y, x, n, p, trait_type, cfg, n_chrom, results, mu, w, residuals,
                       alpha);
}

int saige_fit_null(int device, const uint8_t* bed_body, size_t bpm, size_t n_total,
                   size_t m_total, const uint32_t* sample_idx, size_t n, const double* y,
                   const double* x, size_t p, int trait_type, const saige_fit_null_config* cfg,
                   saige_fit_null_result* result, double* mu_out, double* xvx_inv_xv,
                   uint8_t* marker_class, double* alpha, double* per_marker_vr) {
  clear_error();
  return fit_null_impl(device, bed_body, false, bpm, n_total, m_total, sample_idx, n, y, x, p,
                       trait_type, cfg, result, mu_out, xvx_inv_xv, marker_class, alpha,
                       per_marker_vr);
}

int saige_fit_null_device(int device, const uint8_t* d_bed_body, size_t bpm, size_t n_total,
                          size_t m_total, const uint32_t* sample_idx, size_t n, const double* y,
                          const double* x, size_t p, int trait_type,
                          const saige_fit_null_config* cfg, saige_fit_null_result* result,
                          double* mu_out, double* xvx_inv_xv, uint8_t* marker_class,
                          double* alpha, double* per_marker_vr) {
  clear_error();
  return fit_null_impl(device, d_bed_body, true, bpm, n_total, m_total, sample_idx, n, y, x, p,
                       trait_type, cfg, result, mu_out, xvx_inv_xv, marker_class, alpha,
                       per_marker_vr);
}


int saige_fit_null_multi(int device, const uint8_t* bed_body, size_t bpm, size_t n_total,
                         size_t m_total, const uint32_t* sample_idx, size_t n, const double* y,
                         size_t n_traits, const double* x, size_t p, const int* trait_types,
                         const saige_fit_null_config* cfg, saige_fit_null_result* results,
                         double* mu_out, double* xvx_inv_xv, uint8_t* marker_class,
                         double* alpha, double* per_marker_vr) {
  clear_error();
  return fit_null_multi_impl(device, bed_body, false, bpm, n_total, m_total, sample_idx, n, y,
                             n_traits, x, p, trait_types, cfg, results, mu_out, xvx_inv_xv,
                             marker_class, alpha, per_marker_vr);
}

int saige_fit_null_multi_device(int device, const uint8_t* d_bed_body, size_t bpm,
                                size_t n_total, size_t m_total, const uint32_t* sample_idx,
                                size_t n, const double* y, size_t n_traits, const double* x,
                                size_t p, const int* trait_types,
                                const saige_fit_null_config* cfg, saige_fit_null_result* results,
                                double* mu_out, double* xvx_inv_xv, uint8_t* marker_class,
                                double* alpha, double* per_marker_vr) {
  clear_error();
  return fit_null_multi_impl(device, d_bed_body, true, bpm, n_total, m_total, sample_idx, n, y,
                             n_traits, x, p, trait_types, cfg, results, mu_out, xvx_inv_xv,
                             marker_class, alpha, per_marker_vr);
}

}  // extern "C"
