// Step 2 on the GPU: the single-variant score test with variance ratio and the
// saddle-point approximation.
//
// Reference: ScoreTestEngine::test_marker (saige-core/src/score_test/single_variant.rs:
// 112-257), compute_case_control_af (:260-291), select_variance_ratio
// (glmm/variance_ratio.rs:289-301), spa_binary_fast (spa/fast.rs:207-242) with
// find_root_k1_fast (:53-132), saddle_probability_fast (:135-180) and
// partition_for_fast_spa (:185-204).
//
// Kernel 1 (score_sums): a batch of variants is a (variants x samples) matrix times a
// (samples x features) matrix.  One lane per variant, 64 variants per wave; the lanes of
// a wave walk the same sample index, so the per-sample feature row
//   [A(:, i) (p values), res_i, mu2_i, mu2_i * x_i (p values), mu_i, case_i]
// is wave-uniform and comes from the scalar cache.  The sample range is split over the
// grid; partial sums are combined in a fixed order (deterministic).
// Kernel 2 (score_finalize): one thread per variant turns the sums into S, varT, varT*,
// T, the chi-square(1) p-value and the SPA gate, using
//   g~ = g - X (A g)                       (single_variant.rs:137-144)
//   g~'res      = g'res - (A g)'(X'res)
//   sum mu2 g~^2 = g'M2g - 2 (A g)'(X'M2 g) + (A g)'(X'M2X)(A g)
//   sum mu g~    = g'mu - (A g)'(X'mu)
// which hold for any p x n matrix A (see the layout note on xvx_inv_xv_flat).
// Kernel 3 (spa): one work-group per gated variant runs both safeguarded Newton roots in
// lockstep, recomputing g~_i from the dosage, X and A g in every pass, each pass fusing
// K'(t) and K''(t) (the values the reference's loop needs at the same point), then the
// Lugannani-Rice tail.  Control flow per root is exactly find_root_k1_fast.
#include "common.hpp"
#include "mlp_common.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <type_traits>

namespace saige {

constexpr int kMaxP = 8;
constexpr int kMaxCat = 16;

struct EngineConst {
  double tau_e;
  double vr;
  double cutoff;
  double spa_tol;
  int use_spa;
  int fast;
  int n_cat;
  int has_y;
  double n_cases;  // cases among all samples (the count when no dosage can be missing)
  double xtr[kMaxP];
  double xtm[kMaxP];
  double xm2x[kMaxP * kMaxP];
  double cat_upper[kMaxCat];
  double cat_vr[kMaxCat];
};

struct GateInfo {
  uint32_t var;
  uint32_t pad;
  double q, qinv, pn;
  double sd;       // sqrt(var_t / tau_e) = sqrt(K''(0)): the scale of the saddle point
  double feat[6];  // features of the learned root proposal (root 1; root 2 negates feat[0])
  double ag[kMaxP];
};

// Learned root proposal used as the Newton starting point (mlp.hip).
struct MlpView {
  const float* w;  // fp32 copy of the bf16 weights (exactly bf16-representable values)
  float mean[6], inv_std[6];
  float y_scale;
  int enabled;
};
void mlp_view(const saige_mlp* h, const float** w, float* mean, float* inv_std, float* y_scale);

// Both roots' starting points from the learned proposal; all threads of the block call it
// and all receive the result.  128 threads: thread (r, j) owns hidden unit j of root r.
__device__ void mlp_propose2(const MlpView& mv, const GateInfo& g, float* sh, double* tinit) {
  const int tid = threadIdx.x, r = tid >> 6, j = tid & 63;
  const float* w = mv.w;
  if (tid < 128) {
    float a = w[kMlpB1 + j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) {
      const float f = (float)((i == 0 && r == 1) ? -g.feat[0] : g.feat[i]);
      a = fmaf(w[kMlpW1 + j * kMlpIn + i], (f - mv.mean[i]) * mv.inv_std[i], a);
    }
    sh[r * kMlpH + j] = mlp_gelu(a);
  }
  __syncthreads();
  float h2 = 0.0f;
  if (tid < 128) {
    float a = w[kMlpB2 + j];
    for (int k = 0; k < kMlpH; ++k) a = fmaf(w[kMlpW2 + j * kMlpH + k], sh[r * kMlpH + k], a);
    h2 = w[kMlpW3 + j] * mlp_gelu(a);
  }
  __syncthreads();
  if (tid < 128) sh[r * kMlpH + j] = h2;
  __syncthreads();
  if (tid < 2) {
    float o = w[kMlpB3];
    for (int k = 0; k < kMlpH; ++k) o += sh[tid * kMlpH + k];
    const double t = sinh((double)o * (double)mv.y_scale) / g.sd;
    sh[2 * kMlpH + tid] = isfinite(t) ? (float)t : 0.0f;
  }
  __syncthreads();
  tinit[0] = (double)sh[2 * kMlpH];
  tinit[1] = (double)sh[2 * kMlpH + 1];
  __syncthreads();
}

__device__ __forceinline__ double dnan() { return __longlong_as_double(0x7ff8000000000000ll); }

// ---- dosage loaders: value of sample i of variant v (NaN = missing) ----
struct LoadU8 {
  static constexpr bool kMayMiss = false;
  const uint8_t* q;
  size_t ld;
  double scale;
  int vec;  // rows 16-byte aligned: 16 samples per load
  __device__ __forceinline__ double operator()(size_t v, size_t i) const {
    return (double)q[v * ld + i] * scale;
  }
  __device__ __forceinline__ void load16(size_t v, size_t i, double* g) const {
    expand16(raw16(v, i), g);
  }
  __device__ __forceinline__ uint4 raw16(size_t v, size_t i) const {
    return *reinterpret_cast<const uint4*>(q + v * ld + i);
  }
  __device__ __forceinline__ void expand16(uint4 w, double* g) const {
    const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = (double)((u[k >> 2] >> (8 * (k & 3))) & 0xffu) * scale;
  }
};
struct LoadF64 {
  static constexpr bool kMayMiss = true;
  const double* g;
  size_t ld;
  int vec;
  __device__ __forceinline__ double operator()(size_t v, size_t i) const { return g[v * ld + i]; }
  __device__ __forceinline__ void load16(size_t v, size_t i, double* out) const {
#pragma unroll
    for (int k = 0; k < 16; ++k) out[k] = g[v * ld + i + k];
  }
};
struct LoadPacked {  // plink.rs:170-194, 237-247
  static constexpr bool kMayMiss = true;
  const uint8_t* bed;
  size_t bpm;
  const uint32_t* idx;
  int vec;  // no subset and 4-byte aligned rows: 16 samples per 32-bit load
  __device__ __forceinline__ static double dec(uint32_t c) {
    return c == 0u ? 2.0 : (c == 2u ? 1.0 : (c == 3u ? 0.0 : dnan()));
  }
  __device__ __forceinline__ double operator()(size_t v, size_t i) const {
    size_t s = idx ? idx[i] : i;
    return dec((bed[v * bpm + (s >> 2)] >> (2 * (s & 3))) & 3u);
  }
  __device__ __forceinline__ void load16(size_t v, size_t i, double* g) const {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(bed + v * bpm + (i >> 2));
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = dec((w >> (2 * k)) & 3u);
  }
};

// feature row layout
template <int P> struct Layout {
  static constexpr int NF = 2 * P + 4;  // A(p), res, mu2, mu2*x(p), mu, case
  // sums: g, n, gA(p), g res, g^2 mu2, g mu2 x(p), g mu, g case, n case, g ctrl, and over
  // carriers (g > 0): mu, mu2, mu * mu2 (the moments the learned root proposal reads)
  static constexpr int NS = 2 * P + 12;
};

template <int P, bool MISS>
__device__ __forceinline__ void accumulate(double g, const double* __restrict__ f, double* acc) {
  bool ok = true;
  if (MISS) ok = !isnan(g);
  const double gz = ok ? g : 0.0;
  acc[0] += gz;
  if (MISS) acc[1] += ok ? 1.0 : 0.0;
#pragma unroll
  for (int r = 0; r < P; ++r) acc[2 + r] = fma(gz, f[r], acc[2 + r]);
  acc[2 + P] = fma(gz, f[P], acc[2 + P]);
  acc[3 + P] = fma(gz * gz, f[P + 1], acc[3 + P]);
#pragma unroll
  for (int j = 0; j < P; ++j) acc[4 + P + j] = fma(gz, f[P + 2 + j], acc[4 + P + j]);
  acc[4 + 2 * P] = fma(gz, f[2 * P + 2], acc[4 + 2 * P]);
  const double cs = f[2 * P + 3];
  acc[5 + 2 * P] = fma(gz, cs, acc[5 + 2 * P]);
  if (MISS) acc[6 + 2 * P] += ok ? cs : 0.0;
  acc[7 + 2 * P] = fma(gz, 1.0 - cs, acc[7 + 2 * P]);
  if (gz > 0.0) {
    const double m = f[2 * P + 2], m2 = f[P + 1];
    acc[8 + 2 * P] += m;
    acc[9 + 2 * P] += m2;
    acc[10 + 2 * P] = fma(m, m2, acc[10 + 2 * P]);
  }
}

// Without possible missing values the counts n_valid and n_cases are the engine's
// constants and are filled in by the finalize kernel.
template <int P, class L>
__global__ __launch_bounds__(64) void score_sums_kernel(L load, const double* __restrict__ feat,
                                                        size_t n, size_t nvar, size_t sps,
                                                        double* __restrict__ part,
                                                        size_t nvar_pad) {
  constexpr int NF = Layout<P>::NF, NS = Layout<P>::NS;
  constexpr bool MISS = L::kMayMiss;
  const size_t v = (size_t)blockIdx.x * 64 + threadIdx.x;
  const size_t vv = v < nvar ? v : nvar - 1;
  const size_t i0 = (size_t)blockIdx.y * sps;
  const size_t i1 = min(i0 + sps, n);
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  size_t i = i0;
  if constexpr (std::is_same<L, LoadU8>::value) {
    if (load.vec && i + 16 <= i1) {
      // the uint8 stream: the next two 16-sample segments of the lane's row are in flight
      // while the current one is accumulated
      uint4 w0 = load.raw16(vv, i);
      uint4 w1 = i + 32 <= i1 ? load.raw16(vv, i + 16) : w0;
      for (; i + 16 <= i1; i += 16) {
        const uint4 cur = w0;
        w0 = w1;
        if (i + 48 <= i1) w1 = load.raw16(vv, i + 32);
        double g[16];
        load.expand16(cur, g);
#pragma unroll
        for (int k = 0; k < 16; ++k) accumulate<P, MISS>(g[k], feat + (i + k) * NF, acc);
      }
    }
  } else if (load.vec) {
    for (; i + 16 <= i1; i += 16) {
      double g[16];
      load.load16(vv, i, g);
#pragma unroll
      for (int k = 0; k < 16; ++k) accumulate<P, MISS>(g[k], feat + (i + k) * NF, acc);
    }
  }
  for (; i < i1; ++i) accumulate<P, MISS>(load(vv, i), feat + i * NF, acc);
  if (v >= nvar) return;
#pragma unroll
  for (int s = 0; s < NS; ++s) part[((size_t)blockIdx.y * NS + s) * nvar_pad + v] = acc[s];
}

__device__ __forceinline__ double chisq1_cdf(double t) { return t <= 0.0 ? 0.0 : erf(sqrt(0.5 * t)); }
__device__ __forceinline__ double norm_cdf(double z) { return 0.5 * erfc(-z / sqrt(2.0)); }

template <int P>
__global__ void score_finalize_kernel(const double* __restrict__ part, size_t splits,
                                      size_t nvar, size_t nvar_pad, size_t n, EngineConst ec,
                                      int may_miss, saige_variant_result* __restrict__ out,
                                      GateInfo* __restrict__ gate, uint32_t* __restrict__ n_gate) {
  constexpr int NS = Layout<P>::NS;
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvar) return;
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double a = 0.0;
    for (size_t q = 0; q < splits; ++q) a += part[(q * NS + k) * nvar_pad + v];
    s[k] = a;
  }
  if (!may_miss) {
    s[1] = (double)n;
    s[6 + 2 * P] = ec.n_cases;
  }
  const double sum = s[0], nv = s[1];
  saige_variant_result r;
  r.n = (uint64_t)nv;
  r.ac = sum;
  r.af = sum / (2.0 * nv);
  r.mac = fmin(sum, 2.0 * nv - sum);
  const bool missing = nv < (double)n;  // a NaN dosage makes A g, and so g~, NaN
  double ag[P];
#pragma unroll
  for (int j = 0; j < P; ++j) ag[j] = missing ? dnan() : s[2 + j];
  double gres = s[2 + P];
  double q2 = s[3 + P];
  double gm = s[4 + 2 * P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    gres -= ag[j] * ec.xtr[j];
    q2 -= 2.0 * ag[j] * s[4 + P + j];
    gm -= ag[j] * ec.xtm[j];
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < P; ++k) t += ec.xm2x[j * P + k] * ag[k];
    q2 += ag[j] * t;
  }
  const double score = gres / ec.tau_e;
  double vr = ec.vr;
  for (int c = 0; c < ec.n_cat; ++c)
    if (r.mac <= ec.cat_upper[c]) { vr = ec.cat_vr[c]; break; }
  const double vt = q2 * ec.tau_e;
  const double vts = vt * vr;
  const double tstat = vts > 1e-30 ? score * score / vts : 0.0;
  const double pn = 1.0 - chisq1_cdf(tstat);
  r.score = score;
  r.var_t = vt;
  r.var_t_star = vts;
  r.tstat = tstat;
  r.pvalue = pn;
  r.pvalue_noadj = pn;
  r.is_spa = 0;
  r.spa_converged = 1;
  r.spa_iters1 = 0;
  r.spa_iters2 = 0;
  r.spa_root1 = 0.0;
  r.spa_root2 = 0.0;
  {
    const double sd = sqrt(vt / ec.tau_e);
    const double f[6] = {vts > 0.0 ? score / sqrt(vts) : 0.0, log(fmax(vt, 1e-300)),
                         log1p(fmax(r.mac, 0.0)), log(s[8 + 2 * P] + 1e-8),
                         log(s[9 + 2 * P] + 1e-8), log(s[10 + 2 * P] + 1e-8)};
    for (int j = 0; j < 6; ++j) r.spa_features[j] = (float)f[j];
    (void)sd;
  }
  r.variance_ratio = vr;
  r.beta = vts > 1e-30 ? score / vts : 0.0;
  r.se_beta = vts > 1e-30 ? 1.0 / sqrt(vts) : __longlong_as_double(0x7ff0000000000000ll);
  if (ec.has_y) {
    const double sc = s[5 + 2 * P], nc = s[6 + 2 * P], sn = s[7 + 2 * P];
    const double nn = nv - nc;
    r.af_cases = nc > 0.0 ? sc / (2.0 * nc) : dnan();
    r.af_controls = nn > 0.0 ? sn / (2.0 * nn) : dnan();
    r.n_cases = (uint64_t)nc;
    r.n_controls = (uint64_t)nn;
  } else {
    r.af_cases = dnan();
    r.af_controls = dnan();
    r.n_cases = 0;
    r.n_controls = 0;
  }
  if (ec.use_spa && pn < ec.cutoff) {  // single_variant.rs:183-207
    const double q = score / sqrt(vr) + gm;
    const double qinv = 2.0 * gm - q;
    uint32_t slot = atomicAdd(n_gate, 1u);
    GateInfo gi;
    gi.var = (uint32_t)v;
    gi.pad = 0;
    gi.q = q;
    gi.qinv = qinv;
    gi.pn = pn;
    gi.sd = sqrt(vt / ec.tau_e);
    for (int j = 0; j < 6; ++j) gi.feat[j] = (double)r.spa_features[j];
#pragma unroll
    for (int j = 0; j < kMaxP; ++j) gi.ag[j] = j < P ? ag[j] : 0.0;
    gate[slot] = gi;
  }
  out[v] = r;
}

// ---------------------------------------------------------------------------------
// SPA
// ---------------------------------------------------------------------------------
constexpr int kSpaBlock = 1024;  // 16 waves: 4 per SIMD hide the fp64 latency of one variant

__device__ __forceinline__ void block_sum4(double* v, double* sh) {
  // sh: 4 * kSpaBlock doubles; deterministic tree
  for (int k = 0; k < 4; ++k) sh[k * kSpaBlock + threadIdx.x] = v[k];
  __syncthreads();
  for (int w = kSpaBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; ++k) sh[k * kSpaBlock + threadIdx.x] += sh[k * kSpaBlock + threadIdx.x + w];
    __syncthreads();
  }
  for (int k = 0; k < 4; ++k) v[k] = sh[k * kSpaBlock];
  __syncthreads();
}

struct RootState {
  double t, k1e, k2e, prev_jump, pend, root;
  int phase;  // 0 init-eval, 1 at tn, 2 at th, 3 done
  int it, converged;
  int spent;  // iterations of an abandoned start (learned proposal that did not converge)
};

// Begin (or restart) the search at t0 with the state find_root_k1_fast starts from.
__device__ __forceinline__ void root_start(RootState& s, double t0) {
  s.t = t0;
  s.pend = t0;
  s.prev_jump = __longlong_as_double(0x7ff0000000000000ll);
  s.it = 0;
  s.k1e = s.k2e = 0.0;
  s.phase = 0;
  s.converged = 0;
  s.root = 0.0;
}

// After an evaluation at the current point t (k1e = K'(t) - q, k2e = K''(t)), take the
// Newton step of find_root_k1_fast (fast.rs:79-105); returns with either a new pending
// evaluation point or the root decided.
__device__ __forceinline__ void root_step(RootState& s, double tol, int max_iter) {
  const double tn = s.t - s.k1e / s.k2e;
  if (isnan(tn) || isinf(tn)) { s.root = s.t; s.converged = 0; s.phase = 3; return; }
  if (fabs(tn - s.t) < tol) { s.root = tn; s.converged = 1; s.phase = 3; return; }
  if (s.it == max_iter) { s.root = tn; s.converged = 0; s.phase = 3; return; }
  s.pend = tn;
  s.phase = 1;
}

__device__ __forceinline__ void root_after_eval(RootState& s, double nk1, double nk2, double tol,
                                                int max_iter) {
  if (s.phase == 0) {
    s.k1e = nk1;
    s.k2e = nk2;
    s.it = 1;
    root_step(s, tol, max_iter);
    return;
  }
  if (s.phase == 1) {  // evaluated at tn (fast.rs:107-124)
    const double tn = s.pend;
    if (s.k1e * nk1 < 0.0) {
      if (fabs(tn - s.t) > (s.prev_jump - tol)) {
        const double dir = (nk1 - s.k1e > 0.0) ? 1.0 : -1.0;
        const double th = s.t + dir * s.prev_jump / 2.0;
        s.prev_jump /= 2.0;
        s.pend = th;
        s.phase = 2;
        return;
      }
      s.prev_jump = fabs(tn - s.t);
    }
    s.t = tn;
    s.k1e = nk1;
    s.k2e = nk2;
    s.it += 1;
    root_step(s, tol, max_iter);
    return;
  }
  // phase 2: evaluated at the halved point
  s.t = s.pend;
  s.k1e = nk1;
  s.k2e = nk2;
  s.it += 1;
  root_step(s, tol, max_iter);
}

// g~_i for sample i of the gated variant: g_i - sum_j x_ij (A g)_j (dense.rs:102-113 order).
template <class L, int P>
struct GTilde {
  L load;
  const double* x;  // n x P col-major
  size_t n;
  __device__ __forceinline__ double operator()(size_t v, size_t i, const double* ag) const {
    double xa = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) xa += x[i + j * n] * ag[j];
    return load(v, i) - xa;
  }
};

// g~ for an intercept-only model from uint8 dosages: g~ = q * scale - (A g)_0 depends only
// on the dosage byte, so every exp of a pass is one of 256 table entries (the same values
// the per-sample evaluation would produce).
struct GTildeU8Tab {
  LoadU8 load;
  __device__ __forceinline__ uint32_t byte(size_t v, size_t i) const { return load.q[v * load.ld + i]; }
};

template <class G, bool TAB>
__global__ __launch_bounds__(kSpaBlock) void spa_kernel(G gt, const double* __restrict__ mu,
                                                        size_t n, const GateInfo* __restrict__ gate,
                                                        const uint32_t* __restrict__ n_gate,
                                                        double tol, int fast, MlpView mv,
                                                        saige_variant_result* __restrict__ out,
                                                        saige_spa_result* __restrict__ single) {
  const uint32_t slot = blockIdx.x;
  if (slot >= *n_gate) return;
  __shared__ double sh[4 * kSpaBlock];
  __shared__ GateInfo g;
  __shared__ double tg[TAB ? 256 : 1], te0[TAB ? 256 : 1], te1[TAB ? 256 : 1];
  if (threadIdx.x == 0) g = gate[slot];
  __syncthreads();
  const size_t v = g.var;
  const double q[2] = {g.q, g.qinv};
  const double thr = fast ? 1e-10 : -1.0;  // fast: |g~| > 1e-10 enters K (fast.rs:191-200)
  if constexpr (TAB) {
    for (int b = threadIdx.x; b < 256; b += kSpaBlock) tg[b] = (double)b * gt.load.scale - g.ag[0];
    __syncthreads();
  }
  auto gval = [&](size_t i) -> double {
    if constexpr (TAB) return tg[gt.byte(v, i)];
    else return gt(v, i, g.ag);
  };
  // setup pass: g_pos, g_neg over all samples; NA moments over the excluded samples
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t i = threadIdx.x; i < n; i += kSpaBlock) {
    const double gi = gval(i);
    const double m = mu[i];
    if (gi > 0.0) a[0] += gi;
    if (gi < 0.0) a[1] += gi;
    if (!(fabs(gi) > thr)) {
      a[2] += m * gi;
      a[3] += m * (1.0 - m) * gi * gi;
    }
  }
  block_sum4(a, sh);
  const double gpos = a[0], gneg = a[1], na_mu = a[2], na_sigma = a[3];
  // starting points: 0 as in fast.rs:217-218, or the learned proposal
  double tinit[2] = {0.0, 0.0};
  if (mv.enabled) mlp_propose2(mv, g, reinterpret_cast<float*>(sh), tinit);
  RootState st[2];
  for (int r = 0; r < 2; ++r) {
    root_start(st[r], tinit[r]);
    st[r].spent = 0;
    if (q[r] >= gpos || q[r] <= gneg) {  // fast.rs:67-73: root at infinity
      st[r].root = __longlong_as_double(0x7ff0000000000000ll);
      st[r].converged = 1;
      st[r].phase = 3;
    }
  }
  const int max_iter = 1000;
  // Each pass evaluates K'(t) - q and K''(t) at both roots' pending points.  A search from
  // a learned proposal that fails is restarted once from 0, the plain method's start, so
  // the proposal can only add roots, never lose one the plain method finds.
  for (int guard = 0; guard < 8 * max_iter + 16; ++guard) {
    if (st[0].phase == 3 && st[1].phase == 3) break;
    const double t0 = st[0].pend, t1 = st[1].pend;
    if constexpr (TAB) {
      for (int b = threadIdx.x; b < 256; b += kSpaBlock) {
        te0[b] = exp(-tg[b] * t0);
        te1[b] = exp(-tg[b] * t1);
      }
      __syncthreads();
    }
    double b[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = threadIdx.x; i < n; i += kSpaBlock) {
      double gi, e0, e1;
      if constexpr (TAB) {
        const uint32_t by = gt.byte(v, i);
        gi = tg[by];
        e0 = te0[by];
        e1 = te1[by];
      } else {
        gi = gt(v, i, g.ag);
        e0 = exp(-gi * t0);
        e1 = exp(-gi * t1);
      }
      if (!(fabs(gi) > thr)) continue;
      const double m = mu[i];
      const double d0 = (1.0 - m) * e0 + m;
      const double d1 = (1.0 - m) * e1 + m;
      // one reciprocal per root: K' term m g / d, K'' term (1 - m) m g^2 e / d^2
      const double r0 = 1.0 / d0, r1 = 1.0 / d1;
      const double mg = m * gi, w2 = (1.0 - m) * mg * gi;
      b[0] += mg * r0;
      b[1] += w2 * e0 * r0 * r0;
      b[2] += mg * r1;
      b[3] += w2 * e1 * r1 * r1;
    }
    block_sum4(b, sh);
    for (int r = 0; r < 2; ++r) {
      if (st[r].phase == 3) continue;
      const double tt = r == 0 ? t0 : t1;
      const double k1 = b[2 * r] + na_mu + na_sigma * tt - q[r];
      const double k2 = b[2 * r + 1] + na_sigma;
      root_after_eval(st[r], k1, k2, tol, max_iter);
      if (st[r].phase == 3 && !st[r].converged && tinit[r] != 0.0 && st[r].spent == 0) {
        st[r].spent = st[r].it;
        root_start(st[r], 0.0);
      }
    }
  }
  const bool both = st[0].converged && st[1].converged;
  double p1 = 0.0, p2 = 0.0;
  bool ok1 = false, ok2 = false;
  if (both) {
    // saddle_probability_fast (fast.rs:135-180): K(zeta), K''(zeta) at both roots
    const double z0 = st[0].root, z1 = st[1].root;
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = threadIdx.x; i < n; i += kSpaBlock) {
      const double gi = gval(i);
      if (!(fabs(gi) > thr)) continue;
      const double m = mu[i];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const double zz = r == 0 ? z0 : z1;
        c[2 * r] += log(1.0 - m + m * exp(gi * zz));
        const double e = exp(-gi * zz);
        const double den = (1.0 - m) * e + m;
        c[2 * r + 1] += (1.0 - m) * m * gi * gi * e / (den * den);
      }
    }
    block_sum4(c, sh);
    for (int r = 0; r < 2; ++r) {
      const double zeta = r == 0 ? z0 : z1;
      const double k0 = c[2 * r] + na_mu * zeta + 0.5 * na_sigma * zeta * zeta;
      const double k2 = c[2 * r + 1] + na_sigma;
      const double temp1 = zeta * q[r] - k0;
      double pv = 0.0;
      bool ok = false;
      if (isfinite(k0) && isfinite(k2) && !(temp1 < 0.0) && !(k2 < 0.0)) {
        const double sg = zeta > 0.0 ? 1.0 : (zeta < 0.0 ? -1.0 : (signbit(zeta) ? -1.0 : 1.0));
        const double w = (isnan(zeta) ? zeta : sg) * sqrt(2.0 * temp1);
        const double vv = zeta * sqrt(k2);
        if (w != 0.0) {
          const double z = w + (1.0 / w) * log(vv / w);
          pv = z > 0.0 ? 1.0 - norm_cdf(z) : -norm_cdf(z);
          ok = true;
        }
      }
      if (r == 0) { p1 = pv; ok1 = ok; } else { p2 = pv; ok2 = ok; }
    }
  }
  if (threadIdx.x == 0) {
    double pval;
    int conv, is_spa;
    if (both) {
      const double a1 = ok1 ? p1 : g.pn / 2.0;
      const double a2 = ok2 ? p2 : g.pn / 2.0;
      pval = fabs(a1) + fabs(a2);
      conv = 1;
      is_spa = (ok1 || ok2) ? 1 : 0;
    } else {
      pval = g.pn;
      conv = 0;
      is_spa = 0;
    }
    if (out) {
      out[v].pvalue = pval;
      out[v].is_spa = is_spa;
      out[v].spa_converged = conv;
      out[v].spa_iters1 = st[0].spent + st[0].it;
      out[v].spa_iters2 = st[1].spent + st[1].it;
      out[v].spa_root1 = st[0].root;
      out[v].spa_root2 = st[1].root;
    }
    if (single) {
      single->pvalue = pval;
      single->converged = conv;
      single->is_spa = is_spa;
      single->iters1 = st[0].spent + st[0].it;
      single->iters2 = st[1].spent + st[1].it;
      single->root1 = st[0].root;
      single->root2 = st[1].root;
    }
  }
}


// ---------------------------------------------------------------------------------
// Batched SPA: the passes of spa_kernel with each gated variant's samples split over S
// work-groups.  A pass kernel writes per-slice partial sums; a step kernel (one thread per
// variant) adds them in slice order and advances both root searches with the same state
// machine (root_after_eval, restart from 0 after a failed proposal).  A batch's few
// hundred gated variants then fill the GPU instead of occupying one work-group each.
// ---------------------------------------------------------------------------------
struct SpaVar {
  RootState st[2];
  double gpos, gneg, na_mu, na_sigma;
  double tinit[2];
  int done;  // both searches finished
  int pad;
};
constexpr int kSpaPassBlock = 256;
constexpr int kSpaMaxPasses = 8 * 1000 + 16;  // the guard of spa_kernel's pass loop

template <int NT>
__device__ __forceinline__ void block_sum4t(double* v, double* sh) {
  for (int k = 0; k < 4; ++k) sh[k * NT + threadIdx.x] = v[k];
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; ++k) sh[k * NT + threadIdx.x] += sh[k * NT + threadIdx.x + w];
    __syncthreads();
  }
  for (int k = 0; k < 4; ++k) v[k] = sh[k * NT];
  __syncthreads();
}

// KIND 0: g_pos, g_neg and the NA moments; 1: K'(t) and K''(t) sums at both pending points;
// 2: K(zeta) and K''(zeta) sums at both roots (variants whose two searches converged).
template <class G, bool TAB, int KIND>
__global__ __launch_bounds__(kSpaPassBlock) void spa_pass_kernel(
    G gt, const double* __restrict__ mu, size_t n, const GateInfo* __restrict__ gate,
    const uint32_t* __restrict__ n_gate, int fast, size_t slice,
    const SpaVar* __restrict__ vars, double* __restrict__ part) {
  const uint32_t slot = blockIdx.x;  // grid: (gated variants, slices)
  if (slot >= *n_gate) return;
  if (KIND == 1 && vars[slot].done) return;
  if (KIND == 2 && !(vars[slot].st[0].converged && vars[slot].st[1].converged)) return;
  __shared__ double sh[4 * kSpaPassBlock];
  __shared__ GateInfo g;
  __shared__ double tg[TAB ? 256 : 1], te0[TAB ? 256 : 1], te1[TAB ? 256 : 1];
  if (threadIdx.x == 0) g = gate[slot];
  __syncthreads();
  const size_t v = g.var;
  const double thr = fast ? 1e-10 : -1.0;  // fast: |g~| > 1e-10 enters K (fast.rs:191-200)
  if constexpr (TAB) {
    for (int b = threadIdx.x; b < 256; b += kSpaPassBlock) tg[b] = (double)b * gt.load.scale - g.ag[0];
  }
  double t0 = 0.0, t1 = 0.0;
  if (KIND == 1) {
    t0 = vars[slot].st[0].pend;
    t1 = vars[slot].st[1].pend;
  } else if (KIND == 2) {
    t0 = vars[slot].st[0].root;
    t1 = vars[slot].st[1].root;
  }
  if constexpr (TAB) {
    if (KIND != 0) {
      __syncthreads();
      for (int b = threadIdx.x; b < 256; b += kSpaPassBlock) {
        te0[b] = exp(-tg[b] * t0);
        te1[b] = exp(-tg[b] * t1);
      }
    }
    __syncthreads();
  }
  const size_t i0 = (size_t)blockIdx.y * slice, i1 = min(i0 + slice, n);
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t i = i0 + threadIdx.x; i < i1; i += kSpaPassBlock) {
    double gi, e0 = 1.0, e1 = 1.0;
    if constexpr (TAB) {
      const uint32_t by = gt.byte(v, i);
      gi = tg[by];
      if (KIND != 0) {
        e0 = te0[by];
        e1 = te1[by];
      }
    } else {
      gi = gt(v, i, g.ag);
      if (KIND != 0) {
        e0 = exp(-gi * t0);
        e1 = exp(-gi * t1);
      }
    }
    const double m = mu[i];
    if (KIND == 0) {
      if (gi > 0.0) a[0] += gi;
      if (gi < 0.0) a[1] += gi;
      if (!(fabs(gi) > thr)) {
        a[2] += m * gi;
        a[3] += m * (1.0 - m) * gi * gi;
      }
      continue;
    }
    if (!(fabs(gi) > thr)) continue;
    const double d0 = (1.0 - m) * e0 + m;
    const double d1 = (1.0 - m) * e1 + m;
    const double r0 = 1.0 / d0, r1 = 1.0 / d1;
    if (KIND == 1) {
      const double mg = m * gi, w2 = (1.0 - m) * mg * gi;
      a[0] += mg * r0;
      a[1] += w2 * e0 * r0 * r0;
      a[2] += mg * r1;
      a[3] += w2 * e1 * r1 * r1;
    } else {
      // K at the roots is ln(1 - m + m e^{g zeta}) (binary.rs:148-151), K'' as above
      const double w2 = (1.0 - m) * m * gi * gi;
      a[0] += log(1.0 - m + m * exp(gi * t0));
      a[1] += w2 * e0 * r0 * r0;
      a[2] += log(1.0 - m + m * exp(gi * t1));
      a[3] += w2 * e1 * r1 * r1;
    }
  }
  block_sum4t<kSpaPassBlock>(a, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[((size_t)slot * gridDim.y + blockIdx.y) * 4 + k] = a[k];
}

__device__ __forceinline__ double slice_sum(const double* __restrict__ part, uint32_t slot, int S,
                                            int k) {
  double s = 0.0;
  for (int q = 0; q < S; ++q) s += part[((size_t)slot * S + q) * 4 + k];
  return s;
}

// learned starting points of both searches (one 128-thread group per gated variant)
__global__ __launch_bounds__(128) void spa_propose_kernel(const GateInfo* __restrict__ gate,
                                                          const uint32_t* __restrict__ n_gate,
                                                          MlpView mv, SpaVar* __restrict__ vars) {
  const uint32_t slot = blockIdx.x;
  if (slot >= *n_gate) return;
  __shared__ GateInfo g;
  __shared__ float sh[2 * kMlpH + 2];
  if (threadIdx.x == 0) g = gate[slot];
  __syncthreads();
  double tinit[2];
  mlp_propose2(mv, g, sh, tinit);
  if (threadIdx.x == 0) {
    vars[slot].tinit[0] = tinit[0];
    vars[slot].tinit[1] = tinit[1];
  }
}

__global__ void spa_setup_step(const GateInfo* __restrict__ gate, const uint32_t* __restrict__ n_gate,
                               int S, const double* __restrict__ part, int use_tinit,
                               SpaVar* __restrict__ vars) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= *n_gate) return;
  SpaVar V = vars[slot];
  if (!use_tinit) V.tinit[0] = V.tinit[1] = 0.0;  // fast.rs:217-218
  V.gpos = slice_sum(part, slot, S, 0);
  V.gneg = slice_sum(part, slot, S, 1);
  V.na_mu = slice_sum(part, slot, S, 2);
  V.na_sigma = slice_sum(part, slot, S, 3);
  const double q[2] = {gate[slot].q, gate[slot].qinv};
  for (int r = 0; r < 2; ++r) {
    root_start(V.st[r], V.tinit[r]);
    V.st[r].spent = 0;
    if (q[r] >= V.gpos || q[r] <= V.gneg) {  // fast.rs:67-73: root at infinity
      V.st[r].root = __longlong_as_double(0x7ff0000000000000ll);
      V.st[r].converged = 1;
      V.st[r].phase = 3;
    }
  }
  V.done = (V.st[0].phase == 3 && V.st[1].phase == 3) ? 1 : 0;
  vars[slot] = V;
}

// active[pass & 1] counts the variants still searching after this pass; the other counter
// is cleared for the next pass.
__global__ void spa_eval_step(const GateInfo* __restrict__ gate, const uint32_t* __restrict__ n_gate,
                              int S, const double* __restrict__ part, double tol, int pass,
                              SpaVar* __restrict__ vars, uint32_t* __restrict__ active) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot == 0) active[(pass + 1) & 1] = 0u;
  if (slot >= *n_gate) return;
  SpaVar V = vars[slot];
  if (V.done) return;
  const double q[2] = {gate[slot].q, gate[slot].qinv};
  const int max_iter = 1000;
  for (int r = 0; r < 2; ++r) {
    if (V.st[r].phase == 3) continue;
    const double tt = V.st[r].pend;
    const double k1 = slice_sum(part, slot, S, 2 * r) + V.na_mu + V.na_sigma * tt - q[r];
    const double k2 = slice_sum(part, slot, S, 2 * r + 1) + V.na_sigma;
    root_after_eval(V.st[r], k1, k2, tol, max_iter);
    if (V.st[r].phase == 3 && !V.st[r].converged && V.tinit[r] != 0.0 && V.st[r].spent == 0) {
      V.st[r].spent = V.st[r].it;
      root_start(V.st[r], 0.0);
    }
  }
  V.done = (V.st[0].phase == 3 && V.st[1].phase == 3) ? 1 : 0;
  vars[slot] = V;
  if (!V.done) atomicAdd(&active[pass & 1], 1u);
}

// saddle-point probabilities (fast.rs:135-180 and :220-241) and the outputs
__global__ void spa_final_step(const GateInfo* __restrict__ gate, const uint32_t* __restrict__ n_gate,
                               int S, const double* __restrict__ part,
                               const SpaVar* __restrict__ vars,
                               saige_variant_result* __restrict__ out) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= *n_gate) return;
  const SpaVar V = vars[slot];
  const GateInfo g = gate[slot];
  const double q[2] = {g.q, g.qinv};
  const bool both = V.st[0].converged && V.st[1].converged;
  double p[2] = {0.0, 0.0};
  bool ok[2] = {false, false};
  if (both) {
    for (int r = 0; r < 2; ++r) {
      const double zeta = V.st[r].root;
      const double k0 = slice_sum(part, slot, S, 2 * r) + V.na_mu * zeta + 0.5 * V.na_sigma * zeta * zeta;
      const double k2 = slice_sum(part, slot, S, 2 * r + 1) + V.na_sigma;
      const double temp1 = zeta * q[r] - k0;
      if (isfinite(k0) && isfinite(k2) && !(temp1 < 0.0) && !(k2 < 0.0)) {
        const double sg = zeta > 0.0 ? 1.0 : (zeta < 0.0 ? -1.0 : (signbit(zeta) ? -1.0 : 1.0));
        const double w = (isnan(zeta) ? zeta : sg) * sqrt(2.0 * temp1);
        const double vv = zeta * sqrt(k2);
        if (w != 0.0) {
          const double z = w + (1.0 / w) * log(vv / w);
          p[r] = z > 0.0 ? 1.0 - norm_cdf(z) : -norm_cdf(z);
          ok[r] = true;
        }
      }
    }
  }
  double pval;
  int conv, is_spa;
  if (both) {
    const double a1 = ok[0] ? p[0] : g.pn / 2.0;
    const double a2 = ok[1] ? p[1] : g.pn / 2.0;
    pval = fabs(a1) + fabs(a2);
    conv = 1;
    is_spa = (ok[0] || ok[1]) ? 1 : 0;
  } else {
    pval = g.pn;
    conv = 0;
    is_spa = 0;
  }
  saige_variant_result& o = out[g.var];
  o.pvalue = pval;
  o.is_spa = is_spa;
  o.spa_converged = conv;
  o.spa_iters1 = V.st[0].spent + V.st[0].it;
  o.spa_iters2 = V.st[1].spent + V.st[1].it;
  o.spa_root1 = V.st[0].root;
  o.spa_root2 = V.st[1].root;
}

}  // namespace saige

using namespace saige;

struct saige_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  size_t n = 0, p = 0;
  EngineConst ec{};
  bool intercept_only = false;
  const saige_mlp* mlp = nullptr;
  DevBuf feat, mu_d, x_d, idx, part, res, gate, ngate, in;
  DevBuf spa_vars, spa_part, spa_active;  // batched SPA state
  // split counts of the sums kernel chosen by timing, per (loader, covariates, variant blocks)
  std::map<std::tuple<int, int, size_t>, size_t> tuned_splits;
  MlpView mlp_view() const {
    MlpView v{};
    if (mlp) {
      saige::mlp_view(mlp, &v.w, v.mean, v.inv_std, &v.y_scale);
      v.enabled = 1;
    }
    return v;
  }
};

namespace saige {

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

// Sum the per-split partials of the sums kernel in split order (one thread per variant and
// sum), so the finalize kernel reads one row of sums per variant.
__global__ void sums_reduce_kernel(const double* __restrict__ part, size_t splits, int ns,
                                   size_t nvar_pad, double* __restrict__ out) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (v >= nvar_pad) return;
  double a = 0.0;
  for (size_t q = 0; q < splits; ++q) a += part[(q * ns + k) * nvar_pad + v];
  out[(size_t)k * nvar_pad + v] = a;
}

template <class L> constexpr int loader_id() {
  return std::is_same<L, LoadU8>::value ? 1 : (std::is_same<L, LoadF64>::value ? 2 : 3);
}

// Tile-shape autotuner of the sums kernel: the first batch of a shape times the kernel and
// the reduction of its partials at 16 to 256 sample splits (every candidate computes the
// complete, correct sums; the run that follows recomputes them with the choice) and the
// fastest split count serves the later batches of that shape.  Shapes with fewer than 64 * 1024 samples use the default split count.
template <int P, class L>
static int sums_splits(saige_engine* e, L load, size_t nvar, size_t nvar_pad, size_t* out) {
  constexpr int NS = Layout<P>::NS;
  const size_t n = e->n;
  const size_t vblocks = ceil_div(nvar, 64);
  auto norm = [&](size_t sp) {
    const size_t sps = round_up(ceil_div(n, std::max<size_t>(sp, 1)), 16);
    return ceil_div(n, sps);
  };
  const size_t dflt = norm(std::max<size_t>(1, std::min<size_t>(ceil_div(2048, vblocks),
                                                                 ceil_div(n, 512))));
  const auto key = std::make_tuple(loader_id<L>(), P, vblocks);
  auto it = e->tuned_splits.find(key);
  if (it != e->tuned_splits.end()) {
    *out = it->second;
    return SAIGE_OK;
  }
  if (n < 65536) {
    e->tuned_splits[key] = dflt;
    *out = dflt;
    return SAIGE_OK;
  }
  hipStream_t s = e->stream;
  size_t best = dflt;
  float best_ms = 1e30f;
  hipEvent_t a, b;
  SAIGE_HIP(hipEventCreate(&a));
  SAIGE_HIP(hipEventCreate(&b));
  for (size_t cand : {size_t(16), size_t(32), size_t(64), size_t(128), size_t(256)}) {
    const size_t sp = norm(cand);
    const size_t sps = round_up(ceil_div(n, sp), 16);
    SAIGE_TRY(e->part.ensure((sp + 1) * NS * nvar_pad * 8));
    SAIGE_HIP(hipEventRecord(a, s));
    score_sums_kernel<P, L><<<dim3((unsigned)vblocks, (unsigned)sp), 64, 0, s>>>(
        load, e->feat.as<double>(), n, nvar, sps, e->part.as<double>(), nvar_pad);
    sums_reduce_kernel<<<dim3((unsigned)ceil_div(nvar_pad, 256), (unsigned)NS), 256, 0, s>>>(
        e->part.as<double>(), sp, NS, nvar_pad, e->part.as<double>() + sp * NS * nvar_pad);
    SAIGE_HIP(hipEventRecord(b, s));
    SAIGE_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SAIGE_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best_ms) {
      best_ms = ms;
      best = sp;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  e->tuned_splits[key] = best;
  *out = best;
  return SAIGE_OK;
}

// SPA of a batch's gated variants with S work-groups per variant and pass (about 1,024
// work-groups per pass, slices of at least 2,048 samples); the host reads the number of
// variants still searching after every fourth pass.
template <class G, bool TAB>
static int spa_batched(saige_engine* e, G gt, saige_variant_result* d_out) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  uint32_t ng = 0;
  SAIGE_HIP(hipMemcpyAsync(&ng, e->ngate.p, 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  if (ng == 0) return SAIGE_OK;
  size_t S = std::max<size_t>(1, std::min<size_t>(ceil_div(1024, ng), ceil_div(n, 2048)));
  S = std::min<size_t>(S, 64);
  const size_t slice = round_up(ceil_div(n, S), 256);
  S = ceil_div(n, slice);
  SAIGE_TRY(e->spa_vars.ensure(ng * sizeof(SpaVar)));
  SAIGE_TRY(e->spa_part.ensure(ng * S * 4 * 8));
  SAIGE_TRY(e->spa_active.ensure(64));
  uint32_t* active = e->spa_active.as<uint32_t>();
  SpaVar* vars = e->spa_vars.as<SpaVar>();
  double* part = e->spa_part.as<double>();
  const GateInfo* gate = e->gate.as<GateInfo>();
  const uint32_t* ngp = e->ngate.as<uint32_t>();
  const double* mu = e->mu_d.as<double>();
  const MlpView mv = e->mlp_view();
  const dim3 grid(ng, (unsigned)S);
  const unsigned sb = (unsigned)ceil_div(ng, 256);
  const int fast = e->ec.fast, iS = (int)S;
  zero_u32_kernel<<<1, 1, 0, s>>>(active);
  zero_u32_kernel<<<1, 1, 0, s>>>(active + 1);
  if (mv.enabled) spa_propose_kernel<<<ng, 128, 0, s>>>(gate, ngp, mv, vars);
  spa_pass_kernel<G, TAB, 0><<<grid, kSpaPassBlock, 0, s>>>(gt, mu, n, gate, ngp, fast, slice,
                                                             vars, part);
  spa_setup_step<<<sb, 256, 0, s>>>(gate, ngp, iS, part, mv.enabled, vars);
  for (int pass = 0; pass < kSpaMaxPasses; ++pass) {
    spa_pass_kernel<G, TAB, 1><<<grid, kSpaPassBlock, 0, s>>>(gt, mu, n, gate, ngp, fast, slice,
                                                               vars, part);
    spa_eval_step<<<sb, 256, 0, s>>>(gate, ngp, iS, part, e->ec.spa_tol, pass, vars, active);
    if (pass % 4 == 3) {
      uint32_t cnt = 0;
      SAIGE_HIP(hipMemcpyAsync(&cnt, active + (pass & 1), 4, hipMemcpyDeviceToHost, s));
      SAIGE_HIP(hipStreamSynchronize(s));
      if (cnt == 0) break;
    }
  }
  spa_pass_kernel<G, TAB, 2><<<grid, kSpaPassBlock, 0, s>>>(gt, mu, n, gate, ngp, fast, slice,
                                                             vars, part);
  spa_final_step<<<sb, 256, 0, s>>>(gate, ngp, iS, part, vars, d_out);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

template <int P, class L>
static int run_batch_p(saige_engine* e, L load, size_t nvar, saige_variant_result* d_out) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  constexpr int NS = Layout<P>::NS;
  const size_t vblocks = ceil_div(nvar, 64);
  const size_t nvar_pad = round_up(nvar, 64);
  size_t splits = 1;
  SAIGE_TRY((sums_splits<P, L>(e, load, nvar, nvar_pad, &splits)));
  const size_t sps = round_up(ceil_div(n, splits), 16);
  SAIGE_TRY(e->part.ensure((splits + 1) * NS * nvar_pad * 8));
  SAIGE_TRY(e->gate.ensure(nvar * sizeof(GateInfo)));
  SAIGE_TRY(e->ngate.ensure(64));
  zero_u32_kernel<<<1, 1, 0, s>>>(e->ngate.as<uint32_t>());  // a kernel, not a memset node
  score_sums_kernel<P, L><<<dim3((unsigned)vblocks, (unsigned)splits), 64, 0, s>>>(
      load, e->feat.as<double>(), n, nvar, sps, e->part.as<double>(), nvar_pad);
  double* sums = e->part.as<double>() + splits * NS * nvar_pad;
  sums_reduce_kernel<<<dim3((unsigned)ceil_div(nvar_pad, 256), (unsigned)NS), 256, 0, s>>>(
      e->part.as<double>(), splits, NS, nvar_pad, sums);
  score_finalize_kernel<P><<<(unsigned)ceil_div(nvar, 128), 128, 0, s>>>(
      sums, 1, nvar, nvar_pad, n, e->ec, L::kMayMiss ? 1 : 0, d_out,
      e->gate.as<GateInfo>(), e->ngate.as<uint32_t>());
  SAIGE_HIP(hipGetLastError());
  if (e->ec.use_spa) {
    bool tab = false;
    if constexpr (std::is_same<L, LoadU8>::value) tab = (P == 1 && e->intercept_only);
    if (tab) {
      if constexpr (std::is_same<L, LoadU8>::value)
        SAIGE_TRY((spa_batched<GTildeU8Tab, true>(e, GTildeU8Tab{load}, d_out)));
    } else {
      SAIGE_TRY((spa_batched<GTilde<L, P>, false>(e, GTilde<L, P>{load, e->x_d.as<double>(), n},
                                                   d_out)));
    }
  }
  return SAIGE_OK;
}

template <class L>
static int run_batch(saige_engine* e, L load, size_t nvar, saige_variant_result* d_out) {
  switch (e->p) {
    case 1: return run_batch_p<1>(e, load, nvar, d_out);
    case 2: return run_batch_p<2>(e, load, nvar, d_out);
    case 3: return run_batch_p<3>(e, load, nvar, d_out);
    case 4: return run_batch_p<4>(e, load, nvar, d_out);
    case 5: return run_batch_p<5>(e, load, nvar, d_out);
    case 6: return run_batch_p<6>(e, load, nvar, d_out);
    case 7: return run_batch_p<7>(e, load, nvar, d_out);
    case 8: return run_batch_p<8>(e, load, nvar, d_out);
  }
  set_error("p = %zu covariates not supported (1..8)", e->p);
  return SAIGE_ERR_ARG;
}

constexpr size_t kHostBatch = 4096;

}  // namespace saige

extern "C" {

int saige_engine_create(int device, const double* mu, const double* mu2, const double* residuals,
                        double tau_e, double tau_g, const double* a_flat, const double* x,
                        size_t n, size_t p, double variance_ratio, const double* cat_upper,
                        const double* cat_vr, size_t n_cat, int use_spa, int use_fast_spa,
                        double spa_tol, double spa_pval_cutoff, const double* y,
                        saige_engine** out) {
  clear_error();
  SAIGE_ARG(out && mu && mu2 && residuals && a_flat && x, "null pointer");
  *out = nullptr;
  SAIGE_ARG(n > 0, "n > 0");
  SAIGE_ARG(p >= 1 && p <= (size_t)kMaxP, "1 <= p <= 8");
  SAIGE_ARG(n_cat <= (size_t)kMaxCat, "at most 16 categorical ratios");
  SAIGE_ARG(n_cat == 0 || (cat_upper && cat_vr), "categorical arrays");
  (void)tau_g;
  SAIGE_TRY(use_device(device));
  saige_engine* e = new saige_engine();
  e->device = device;
  e->n = n;
  e->p = p;
  if (hipStreamCreate(&e->stream) != hipSuccess) {
    delete e;
    set_error("hipStreamCreate failed");
    return SAIGE_ERR_DEVICE;
  }
  EngineConst& ec = e->ec;
  ec.tau_e = tau_e;
  ec.vr = variance_ratio;
  ec.cutoff = spa_pval_cutoff;
  ec.spa_tol = spa_tol;
  ec.use_spa = use_spa ? 1 : 0;
  ec.fast = use_fast_spa ? 1 : 0;
  ec.n_cat = (int)n_cat;
  ec.has_y = y ? 1 : 0;
  ec.n_cases = 0.0;
  if (y)
    for (size_t i = 0; i < n; ++i) ec.n_cases += (y[i] > 0.5) ? 1.0 : 0.0;
  e->intercept_only = (p == 1);
  for (size_t i = 0; i < n && e->intercept_only; ++i) e->intercept_only = (x[i] == 1.0);
  for (size_t c = 0; c < n_cat; ++c) {
    ec.cat_upper[c] = cat_upper[c];
    ec.cat_vr[c] = cat_vr[c];
  }
  // engine constants X'res, X'mu, X'M2X
  for (size_t j = 0; j < p; ++j) {
    double a = 0.0, b = 0.0;
    for (size_t i = 0; i < n; ++i) {
      a += x[j * n + i] * residuals[i];
      b += x[j * n + i] * mu[i];
    }
    ec.xtr[j] = a;
    ec.xtm[j] = b;
    for (size_t k = 0; k < p; ++k) {
      double c = 0.0;
      for (size_t i = 0; i < n; ++i) c += x[j * n + i] * mu2[i] * x[k * n + i];
      ec.xm2x[j * p + k] = c;
    }
  }
  const size_t nf = 2 * p + 4;
  std::vector<double> feat(n * nf);
  for (size_t i = 0; i < n; ++i) {
    double* f = &feat[i * nf];
    for (size_t r = 0; r < p; ++r) f[r] = a_flat[i * p + r];  // col-major p x n view
    f[p] = residuals[i];
    f[p + 1] = mu2[i];
    for (size_t j = 0; j < p; ++j) f[p + 2 + j] = mu2[i] * x[j * n + i];
    f[2 * p + 2] = mu[i];
    f[2 * p + 3] = (y && y[i] > 0.5) ? 1.0 : 0.0;
  }
  int rc = SAIGE_OK;
  auto fail = [&](int code) {
    saige_engine_destroy(e);
    return code;
  };
  if ((rc = e->feat.ensure(feat.size() * 8)) != SAIGE_OK) return fail(rc);
  if ((rc = e->mu_d.ensure(n * 8)) != SAIGE_OK) return fail(rc);
  if ((rc = e->x_d.ensure(n * p * 8)) != SAIGE_OK) return fail(rc);
  if (hipMemcpy(e->feat.p, feat.data(), feat.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->mu_d.p, mu, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->x_d.p, x, n * p * 8, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("engine upload failed");
    return fail(SAIGE_ERR_DEVICE);
  }
  *out = e;
  return SAIGE_OK;
}

void saige_engine_destroy(saige_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) {
    (void)hipStreamSynchronize(e->stream);
    (void)hipStreamDestroy(e->stream);
  }
  delete e;
}

void* saige_engine_stream(saige_engine* e) { return e ? (void*)e->stream : nullptr; }

int saige_test_markers_u8_device(saige_engine* e, const uint8_t* d_q, size_t ldq, size_t nvar,
                                 double scale, saige_variant_result* d_out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || (d_q && d_out)), "null pointer");
  SAIGE_ARG(ldq >= e->n, "ldq < n");
  if (nvar == 0) return SAIGE_OK;
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  const int vec = (ldq % 16 == 0) && ((uintptr_t)d_q % 16 == 0);
  return run_batch(e, LoadU8{d_q, ldq, scale, vec}, nvar, d_out);
}

int saige_test_markers_u8(saige_engine* e, const uint8_t* q, size_t ldq, size_t nvar, double scale,
                          saige_variant_result* out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || (q && out)), "null pointer");
  SAIGE_ARG(ldq >= e->n, "ldq < n");
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  for (size_t v0 = 0; v0 < nvar; v0 += kHostBatch) {
    size_t nb = std::min(kHostBatch, nvar - v0);
    SAIGE_TRY(e->in.ensure(nb * e->n + nb * sizeof(saige_variant_result) + 256));
    uint8_t* dq = e->in.as<uint8_t>();
    saige_variant_result* dr =
        reinterpret_cast<saige_variant_result*>(e->in.as<uint8_t>() + round_up(nb * e->n, 256));
    SAIGE_HIP(hipMemcpy2DAsync(dq, e->n, q + v0 * ldq, ldq, e->n, nb, hipMemcpyHostToDevice, s));
    SAIGE_TRY(run_batch(e, LoadU8{dq, e->n, scale, (int)(e->n % 16 == 0)}, nb, dr));
    SAIGE_HIP(hipMemcpyAsync(out + v0, dr, nb * sizeof(saige_variant_result), hipMemcpyDeviceToHost, s));
    SAIGE_HIP(hipStreamSynchronize(s));
  }
  return SAIGE_OK;
}

int saige_test_markers_f64(saige_engine* e, const double* g, size_t ldg, size_t nvar,
                           saige_variant_result* out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || (g && out)), "null pointer");
  SAIGE_ARG(ldg >= e->n, "ldg < n");
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  const size_t n = e->n;
  for (size_t v0 = 0; v0 < nvar; v0 += kHostBatch) {
    size_t nb = std::min(kHostBatch, nvar - v0);
    SAIGE_TRY(e->in.ensure(nb * n * 8 + nb * sizeof(saige_variant_result) + 256));
    double* dg = e->in.as<double>();
    saige_variant_result* dr =
        reinterpret_cast<saige_variant_result*>(e->in.as<uint8_t>() + round_up(nb * n * 8, 256));
    SAIGE_HIP(hipMemcpy2DAsync(dg, n * 8, g + v0 * ldg, ldg * 8, n * 8, nb, hipMemcpyHostToDevice, s));
    SAIGE_TRY(run_batch(e, LoadF64{dg, n, 1}, nb, dr));
    SAIGE_HIP(hipMemcpyAsync(out + v0, dr, nb * sizeof(saige_variant_result), hipMemcpyDeviceToHost, s));
    SAIGE_HIP(hipStreamSynchronize(s));
  }
  return SAIGE_OK;
}

int saige_test_markers_plink(saige_engine* e, const uint8_t* bed_body, size_t bpm, size_t n_total,
                             const uint32_t* sample_idx, size_t nvar,
                             saige_variant_result* out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || (bed_body && out)), "null pointer");
  SAIGE_ARG(bpm >= (n_total + 3) / 4, "bytes_per_marker < ceil(n_total/4)");
  if (!sample_idx) SAIGE_ARG(e->n == n_total, "n must equal n_total without a subset");
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  if (sample_idx) {
    for (size_t k = 0; k < e->n; ++k)
      SAIGE_ARG(sample_idx[k] < n_total, "sample_idx out of range");
    SAIGE_TRY(e->idx.ensure(e->n * 4));
    SAIGE_HIP(hipMemcpyAsync(e->idx.p, sample_idx, e->n * 4, hipMemcpyHostToDevice, s));
  }
  for (size_t v0 = 0; v0 < nvar; v0 += kHostBatch) {
    size_t nb = std::min(kHostBatch, nvar - v0);
    SAIGE_TRY(e->in.ensure(nb * bpm + nb * sizeof(saige_variant_result) + 256));
    uint8_t* db = e->in.as<uint8_t>();
    saige_variant_result* dr =
        reinterpret_cast<saige_variant_result*>(e->in.as<uint8_t>() + round_up(nb * bpm, 256));
    SAIGE_HIP(hipMemcpyAsync(db, bed_body + v0 * bpm, nb * bpm, hipMemcpyHostToDevice, s));
    SAIGE_TRY(run_batch(e, LoadPacked{db, bpm, sample_idx ? e->idx.as<uint32_t>() : nullptr,
                                       (int)(!sample_idx && bpm % 4 == 0)}, nb, dr));
    SAIGE_HIP(hipMemcpyAsync(out + v0, dr, nb * sizeof(saige_variant_result), hipMemcpyDeviceToHost, s));
    SAIGE_HIP(hipStreamSynchronize(s));
  }
  return SAIGE_OK;
}

int saige_engine_set_root_proposal(saige_engine* e, const saige_mlp* mlp) {
  clear_error();
  SAIGE_ARG(e, "engine");
  std::lock_guard<std::mutex> lk(e->mu);
  e->mlp = mlp;
  return SAIGE_OK;
}

int saige_spa_binary(int device, const double* mu, const double* g, size_t n, double q,
                     double qinv, double pval_noadj, double tol, int fast,
                     saige_spa_result* out) {
  clear_error();
  SAIGE_ARG(mu && g && out && n > 0, "null pointer");
  SAIGE_TRY(use_device(device));
  hipStream_t s;
  SAIGE_HIP(hipStreamCreate(&s));
  int rc = SAIGE_OK;
  {
    DevBuf dmu, dg, dgate, dn, dres;
    rc = dmu.ensure(n * 8);
    if (rc == SAIGE_OK) rc = dg.ensure(n * 8);
    if (rc == SAIGE_OK) rc = dgate.ensure(sizeof(GateInfo));
    if (rc == SAIGE_OK) rc = dn.ensure(8);
    if (rc == SAIGE_OK) rc = dres.ensure(sizeof(saige_spa_result));
    if (rc == SAIGE_OK) {
      GateInfo gi{};
      gi.var = 0;
      gi.q = q;
      gi.qinv = qinv;
      gi.pn = pval_noadj;
      uint32_t one = 1;
      hipError_t err = hipMemcpyAsync(dmu.p, mu, n * 8, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) err = hipMemcpyAsync(dg.p, g, n * 8, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) err = hipMemcpyAsync(dgate.p, &gi, sizeof(gi), hipMemcpyHostToDevice, s);
      if (err == hipSuccess) err = hipMemcpyAsync(dn.p, &one, 4, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) {
        GTilde<LoadF64, 0> gt{LoadF64{dg.as<double>(), n, 0}, nullptr, n};
        spa_kernel<GTilde<LoadF64, 0>, false><<<1, kSpaBlock, 0, s>>>(gt, dmu.as<double>(), n,
                                                               dgate.as<GateInfo>(), dn.as<uint32_t>(),
                                                               tol, fast, MlpView{}, nullptr,
                                                               dres.as<saige_spa_result>());
        err = hipGetLastError();
      }
      if (err == hipSuccess) err = hipMemcpyAsync(out, dres.p, sizeof(saige_spa_result), hipMemcpyDeviceToHost, s);
      if (err == hipSuccess) err = hipStreamSynchronize(s);
      if (err != hipSuccess) {
        set_error("spa: %s", hipGetErrorString(err));
        rc = SAIGE_ERR_DEVICE;
      }
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

}  // extern "C"
