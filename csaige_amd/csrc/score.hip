// Step 2 on the GPU: the single-variant score test with variance ratio and the
// saddle-point approximation.
//
// Reference: ScoreTestEngine::test_marker (saige-core/src/score_test/single_variant.rs:
// 112-257), compute_case_control_af (:260-291), select_variance_ratio
// (glmm/variance_ratio.rs:289-301), spa_binary_fast (spa/fast.rs:207-242) with
// find_root_k1_fast (:53-132), saddle_probability_fast (:135-180) and
// partition_for_fast_spa (:185-204).
//
// Sums (score_sums_*): a batch of variants is a (variants x samples) dosage matrix against
// the per-sample feature rows [A(:, i) (p values), res_i, mu2_i, mu2_i * x_i (p values),
// case_i].  uint8 dosages take score_sums_u8t_kernel (lanes over samples, several variants
// per wave, feature rows shared through LDS) or score_sums_u8_kernel (a lane per variant,
// feature rows wave-uniform from the scalar cache); the shape is picked by timing the first
// batch of each shape.  Other dosage forms take score_sums_kernel.  The sample range is
// split over the grid; the partial sums are added in a fixed order (deterministic).
// Finalize (score_finalize): one thread per variant turns the sums into S, varT, varT*,
// T, the chi-square(1) p-value and the SPA gate, using
//   g~ = g - X (A g)                       (single_variant.rs:137-144)
//   g~'res      = g'res - (A g)'(X'res)
//   sum mu2 g~^2 = g'M2g - 2 (A g)'(X'M2 g) + (A g)'(X'M2X)(A g)
// which hold for any p x n matrix A (see the layout note on xvx_inv_xv_flat).
// SPA of a batch (spa_pass_*, spa_*_step): the gated variants' passes over the samples are
// split over several work-groups each; a setup pass gives m1 = sum mu g~, the support and
// the carrier moments, then every pass evaluates K'(t) and K''(t) of both roots at their
// pending points and a one-thread-per-variant step advances both safeguarded Newton
// searches with exactly the control flow of find_root_k1_fast, until every search ended;
// a last pass gives K and K'' at the roots for the Lugannani-Rice tails.  spa_kernel (one
// work-group per variant, all passes in one launch) serves single saige_spa_binary calls.
#include <cfloat>

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

// The learned proposal only replaces the path of find_root_k1_fast (fast.rs:53-132) where
// that search, started at 0, is known to reach the same root; everywhere else the search
// runs from 0 and its outcome, a root or the divergence that makes the reference report
// the unadjusted p, is the reference's own.  The search from 0 fails only when an iterate
// reaches |t| where exp(-g t) overflows (|t| > 709 / max|g~|, at least ~354 for dosages in
// [0, 2]) or K''(t) underflows to 0, which takes |t| of the same size; on 2 x 10^4 gated
// roots of seeded cohorts (n = 20,000 and 50,000, prevalence 0.5% to 10%) every failure had
// |root| >= 259, and on roots |root| <= 50 the path of the search from 0 never left
// 4.8 |root|.  So a proposal is used only as a start within kProposalMaxRoot of 0, and a
// search from it that fails or ends at a root beyond kProposalMaxRoot is run again from 0.
constexpr double kProposalMaxRoot = 20.0;

__host__ __device__ __forceinline__ double proposal_start(double t) {
  return (isfinite(t) && fabs(t) <= kProposalMaxRoot) ? t : 0.0;
}

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
    sh[2 * kMlpH + tid] = (float)proposal_start(t);
  }
  __syncthreads();
  tinit[0] = (double)sh[2 * kMlpH];
  tinit[1] = (double)sh[2 * kMlpH + 1];
  __syncthreads();
}

__device__ __forceinline__ double dnan() { return __longlong_as_double(0x7ff8000000000000ll); }

// ---- dosage loaders: value of sample i of variant v (NaN = missing) ----
typedef uint32_t u32x4w __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

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
  __device__ __forceinline__ u32x4w raw16w(size_t v, size_t i) const {
    return *reinterpret_cast<const u32x4w*>(q + v * ld + i);
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
// BGEN v1.2 layout-2 probability data (decompressed), one block per variant at a stride of
// ld bytes, zero-filled past each block (read_bits reads 0 past the data, bgen.rs:439-452).
// info[v]: bits per probability (1..32) | 0x100 when the variant is diploid, biallelic and
// unphased, the case parse_layout2_probabilities decodes (bgen.rs:373-430):
//   dosage = P(AB) + 2 (1 - P(AA) - P(AB)),  P = value / (2^bits - 1)  (second allele)
// Other variants read 0 for every sample that is not missing; a missing sample (bit 7 of its
// ploidy byte) reads NaN.  Probabilities are packed LSB first, sample s at bit 2 s bits.
__device__ __forceinline__ uint64_t bgen_bits(const uint8_t* p, size_t bit, uint32_t nbits) {
  uint64_t r = 0;
  for (uint32_t k = 0; k < nbits; ++k) {
    const size_t t = bit + k;
    r |= (uint64_t)((p[t >> 3] >> (t & 7)) & 1u) << k;
  }
  return r;
}

struct LoadBgen {
  static constexpr bool kMayMiss = true;
  const uint8_t* blk;
  size_t ld;
  const uint32_t* info;
  const uint32_t* idx;  // model sample -> sample of the file, or null
  uint32_t n_file;      // samples per block (the block's N)
  int vec;              // always 0: one sample per call
  __device__ __forceinline__ double operator()(size_t v, size_t i) const {
    const uint8_t* b = blk + v * ld;
    const size_t s = idx ? idx[i] : i;
    if (b[8 + s] & 0x80u) return dnan();
    const uint32_t inf = info[v];
    if (!(inf & 0x100u)) return 0.0;
    const uint32_t bits = inf & 0xffu;
    const double maxv = (double)((1ull << bits) - 1ull);
    const uint8_t* pr = b + 10 + n_file;
    uint64_t a, c;
    if (bits == 8) {
      a = pr[2 * s];
      c = pr[2 * s + 1];
    } else {
      a = bgen_bits(pr, 2 * s * bits, bits);
      c = bgen_bits(pr, (2 * s + 1) * bits, bits);
    }
    const double p_aa = (double)a / maxv, p_ab = (double)c / maxv;
    return p_ab + 2.0 * (1.0 - p_aa - p_ab);
  }
  __device__ __forceinline__ void load16(size_t v, size_t i, double* g) const {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = (*this)(v, i + k);
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
  static constexpr int NF = 2 * P + 3;  // A(p), res, mu2, mu2*x(p), case
  // sums: g, n, gA(p), g res, g^2 mu2, g mu2 x(p), g case, n case.  The moments that only
  // SPA'd variants need (sum of mu g~, carrier moments) come from the SPA setup pass.
  static constexpr int NS = 2 * P + 6;
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
  const double cs = f[2 * P + 2];
  acc[4 + 2 * P] = fma(gz, cs, acc[4 + 2 * P]);
  if (MISS) acc[5 + 2 * P] += ok ? cs : 0.0;
}

// Dosage bytes q of one 16-sample segment against the feature rows: the products use the
// integer-valued q (the dosage scale multiplies the sums once at the end), sum q and the
// case-sample sum of q are exact uint32 dot products with the byte masks.
template <int P>
__device__ __forceinline__ void accumulate_u8(u32x4w w, const double* __restrict__ f,
                                              const uint32_t* __restrict__ cmask, double* acc,
                                              uint32_t& sq, uint32_t& sc) {
  constexpr int NF = Layout<P>::NF;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    sq = __builtin_amdgcn_udot4(w[c], 0x01010101u, sq, false);
    sc = __builtin_amdgcn_udot4(w[c], cmask[c], sc, false);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double q = (double)((w[k >> 2] >> (8 * (k & 3))) & 0xffu);
    const double* fk = f + k * NF;
#pragma unroll
    for (int r = 0; r < P; ++r) acc[2 + r] = fma(q, fk[r], acc[2 + r]);
    acc[2 + P] = fma(q, fk[P], acc[2 + P]);
    acc[3 + P] = fma(q * q, fk[P + 1], acc[3 + P]);
#pragma unroll
    for (int j = 0; j < P; ++j) acc[4 + P + j] = fma(q, fk[P + 2 + j], acc[4 + P + j]);
  }
}

// uint8 dosages, VPL variants per lane (64 * VPL per work-group), one sample range per
// grid row; the next two 16-sample segments of each row are in flight while the current
// one is accumulated.  The feature rows and case masks are wave-uniform (scalar loads).
template <int P, int VPL>
__global__ __launch_bounds__(64) void score_sums_u8_kernel(LoadU8 load,
                                                           const double* __restrict__ feat,
                                                           const uint32_t* __restrict__ cmask,
                                                           size_t n, size_t nvar, size_t sps,
                                                           double* __restrict__ part,
                                                           size_t nvar_pad) {
  constexpr int NF = Layout<P>::NF, NS = Layout<P>::NS;
  size_t vv[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    const size_t v = (size_t)blockIdx.x * 64 * VPL + u * 64 + threadIdx.x;
    vv[u] = v < nvar ? v : nvar - 1;
  }
  const size_t i0 = (size_t)blockIdx.y * sps;
  const size_t i1 = min(i0 + sps, n);
  double acc[VPL][NS];
  uint32_t sq[VPL], sc[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    sq[u] = sc[u] = 0u;
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[u][s] = 0.0;
  }
  size_t i = i0;
  if (i + 16 <= i1) {
    u32x4w w0[VPL], w1[VPL];
#pragma unroll
    for (int u = 0; u < VPL; ++u) {
      w0[u] = load.raw16w(vv[u], i);
      w1[u] = load.raw16w(vv[u], i + 32 <= i1 ? i + 16 : i);
    }
    for (; i + 16 <= i1; i += 16) {
      u32x4w cur[VPL];
      const size_t nx = i + 48 <= i1 ? i + 32 : i;  // in bounds; unused at the tail
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        cur[u] = w0[u];
        w0[u] = w1[u];
        w1[u] = load.raw16w(vv[u], nx);
      }
#pragma unroll
      for (int u = 0; u < VPL; ++u)
        accumulate_u8<P>(cur[u], feat + i * NF, cmask + i / 4, acc[u], sq[u], sc[u]);
    }
  }
  for (; i < i1; ++i) {
    const uint32_t cs = (cmask[i >> 2] >> (8 * (i & 3))) & 0xffu;
#pragma unroll
    for (int u = 0; u < VPL; ++u) {
      const uint32_t qb = load.q[vv[u] * load.ld + i];
      const double q = (double)qb;
      const double* f = feat + i * NF;
      sq[u] += qb;
      sc[u] += qb * cs;
#pragma unroll
      for (int r = 0; r < P; ++r) acc[u][2 + r] = fma(q, f[r], acc[u][2 + r]);
      acc[u][2 + P] = fma(q, f[P], acc[u][2 + P]);
      acc[u][3 + P] = fma(q * q, f[P + 1], acc[u][3 + P]);
#pragma unroll
      for (int j = 0; j < P; ++j) acc[u][4 + P + j] = fma(q, f[P + 2 + j], acc[u][4 + P + j]);
    }
  }
  const double sc1 = load.scale, sc2 = load.scale * load.scale;
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    const size_t v = (size_t)blockIdx.x * 64 * VPL + u * 64 + threadIdx.x;
    if (v >= nvar) continue;
    acc[u][0] = (double)sq[u] * sc1;
    acc[u][4 + 2 * P] = (double)sc[u] * sc1;
#pragma unroll
    for (int s = 2; s < 3 + P; ++s) acc[u][s] *= sc1;
    acc[u][3 + P] *= sc2;
#pragma unroll
    for (int j = 0; j < P; ++j) acc[u][4 + P + j] *= sc1;
#pragma unroll
    for (int s = 0; s < NS; ++s) part[((size_t)blockIdx.y * NS + s) * nvar_pad + v] = acc[u][s];
  }
}

// uint8 dosages with lanes over samples: a wave sums VB variants over its work-group's
// sample range in blocks of 256 samples, four consecutive samples per lane (one 32-bit
// dosage load per variant row, 256 contiguous bytes per wave).  The features of a block are
// staged once per work-group in LDS as [feature][sample] (each thread loads one sample's
// row, the next block's row in flight), so the four waves (4 VB variants) share them; every
// lane keeps per-variant partial sums that are added over the lanes at the end.
template <int P> struct SumsT {
  static constexpr int NFU = 2 * P + 2;             // A(p), res, mu2, mu2*x(p)
  static constexpr int VB = P == 1 ? 8 : 4;         // variants per wave
};

template <int P>
__global__ __launch_bounds__(256) void score_sums_u8t_kernel(LoadU8 load,
                                                             const double* __restrict__ feat,
                                                             const uint32_t* __restrict__ cmask,
                                                             size_t n, size_t nvar, size_t sps,
                                                             double* __restrict__ part,
                                                             size_t nvar_pad) {
  constexpr int NF = Layout<P>::NF, NS = Layout<P>::NS;
  constexpr int NFU = SumsT<P>::NFU, VB = SumsT<P>::VB;
  __shared__ __attribute__((aligned(16))) double fl[2][NFU][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // the wave's variant group is wave-uniform: row offsets live in scalar registers
  const size_t vg = ((size_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(wave)) * VB;
  const size_t i0 = (size_t)blockIdx.y * sps, i1 = min(i0 + sps, n);
  size_t rows[VB];
#pragma unroll
  for (int u = 0; u < VB; ++u) rows[u] = (vg + u < nvar ? vg + u : nvar - 1) * load.ld;
  double acc[VB][NFU];
  uint32_t sq[VB], sc[VB];
#pragma unroll
  for (int u = 0; u < VB; ++u) {
    sq[u] = sc[u] = 0u;
#pragma unroll
    for (int s = 0; s < NFU; ++s) acc[u][s] = 0.0;
  }
  double pf[NFU];
  uint32_t wn[VB], cmn = 0u;
  auto fetch = [&](size_t blk) {
#pragma unroll
    for (int j = 0; j < NFU; ++j) pf[j] = feat[(blk + t) * NF + j];
    const size_t s0 = blk + 4 * lane;
    cmn = cmask[s0 / 4];
#pragma unroll
    for (int u = 0; u < VB; ++u) wn[u] = *reinterpret_cast<const uint32_t*>(load.q + rows[u] + s0);
  };
  auto add = [&](int u, double q, const double* f) {
#pragma unroll
    for (int r = 0; r < P; ++r) acc[u][r] = fma(q, f[r], acc[u][r]);
    acc[u][P] = fma(q, f[P], acc[u][P]);
    acc[u][P + 1] = fma(q * q, f[P + 1], acc[u][P + 1]);
#pragma unroll
    for (int j = 0; j < P; ++j) acc[u][P + 2 + j] = fma(q, f[P + 2 + j], acc[u][P + 2 + j]);
  };
  const size_t iend = i0 + ((i1 - i0) & ~size_t(255));
  int buf = 0;
  if (i0 < iend) fetch(i0);
  for (size_t blk = i0; blk < iend; blk += 256, buf ^= 1) {
#pragma unroll
    for (int j = 0; j < NFU; ++j) fl[buf][j][t] = pf[j];
    uint32_t w[VB];
#pragma unroll
    for (int u = 0; u < VB; ++u) w[u] = wn[u];
    const uint32_t cm = cmn;
    __syncthreads();
    if (blk + 256 < iend) fetch(blk + 256);
    double f[4][NFU];
#pragma unroll
    for (int j = 0; j < NFU; ++j) {
      const f64x2 a = *reinterpret_cast<const f64x2*>(&fl[buf][j][4 * lane]);
      const f64x2 b = *reinterpret_cast<const f64x2*>(&fl[buf][j][4 * lane + 2]);
      f[0][j] = a.x;
      f[1][j] = a.y;
      f[2][j] = b.x;
      f[3][j] = b.y;
    }
#pragma unroll
    for (int u = 0; u < VB; ++u) {
      // 256 samples of dosage 0 add nothing: the wave skips them (rare variants)
      if (__ballot(w[u] != 0u) == 0) continue;
      sq[u] = __builtin_amdgcn_udot4(w[u], 0x01010101u, sq[u], false);
      sc[u] = __builtin_amdgcn_udot4(w[u], cm, sc[u], false);
#pragma unroll
      for (int k = 0; k < 4; ++k) add(u, (double)((w[u] >> (8 * k)) & 0xffu), f[k]);
    }
  }
  // the last < 256 samples of the range, one per lane
  for (size_t i = iend + lane; i < i1; i += 64) {
    double f[NFU];
#pragma unroll
    for (int j = 0; j < NFU; ++j) f[j] = feat[i * NF + j];
    const uint32_t cs = (cmask[i >> 2] >> (8 * (i & 3))) & 0xffu;
#pragma unroll
    for (int u = 0; u < VB; ++u) {
      const uint32_t qb = load.q[rows[u] + i];
      sq[u] += qb;
      sc[u] += qb * cs;
      add(u, (double)qb, f);
    }
  }
  // totals over the 64 lanes (butterfly, the same order in every lane)
#pragma unroll
  for (int u = 0; u < VB; ++u) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
      for (int s = 0; s < NFU; ++s) acc[u][s] += __shfl_xor(acc[u][s], m);
      sq[u] += __shfl_xor(sq[u], m);
      sc[u] += __shfl_xor(sc[u], m);
    }
  }
  if (lane == 0) {
    const double sc1 = load.scale, sc2 = load.scale * load.scale;
#pragma unroll
    for (int u = 0; u < VB; ++u) {
      const size_t v = vg + u;
      if (v >= nvar) continue;
      double o[NS];
      o[0] = (double)sq[u] * sc1;
      o[1] = 0.0;
#pragma unroll
      for (int r = 0; r < P + 1; ++r) o[2 + r] = acc[u][r] * sc1;
      o[3 + P] = acc[u][P + 1] * sc2;
#pragma unroll
      for (int j = 0; j < P; ++j) o[4 + P + j] = acc[u][P + 2 + j] * sc1;
      o[4 + 2 * P] = (double)sc[u] * sc1;
      o[5 + 2 * P] = 0.0;
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) part[((size_t)blockIdx.y * NS + s2) * nvar_pad + v] = o[s2];
    }
  }
}

// uint8 dosages on the int8 matrix cores.  The linear sums (dosage times A(p), the
// residuals and mu2 x(p); the allele count; dosage times the case flag) are exact integer
// dot products: every linear feature f is written once per engine as NDIG signed base-128
// digits of f 2^-E_f (|f| 2^-E_f < 1/2, digits in [-64, 64], the expansion exact to
// 2^(E_f - 7 NDIG)); the dosage byte q enters as q - 128 (its top bit flipped) and
// v_mfma_i32_16x16x64_i8 adds byte times digit over 64 samples into int32.  The +128 of
// q = (q - 128) + 128 is 128 times the digit column sums, which a row of ones through the
// same instruction gives.  The digit columns of a feature are then scaled by 2^(E_f - 7 l)
// and added in double, so a sum differs from the exact sum of the products only by the
// expansion's rounding at 2^(E_f - 70) per sample and the last additions in double.  The
// quadratic sum (q^2 mu2) stays in double on the vector units, skipped for dosage words that
// are zero in every lane.
// Digit table `limbs`: [256-sample block][column][256 samples], zero past n; column 0 ones,
// 1 case flag, 2 + NDIG f + l digit l + 1 of linear feature f (A(p), res, mu2 x(p)).
// A wave owns RT tiles of 16 variants; in step j of a 256-sample block lane (r, qd) holds
// bytes 64 j + 16 qd .. + 15 of variant 16 t + r (each load instruction reads 64 contiguous
// bytes of 16 rows), and the digit table's lanes (column r of a tile, the same samples)
// match them.  Per-split sums go to part as from the other sums kernels.
template <int P> struct SumsM {
  static constexpr int NL = 2 * P + 1;      // linear features
  static constexpr int NDIG = 10;           // base-128 digits per feature
  static constexpr int NC = 2 + NDIG * NL;  // ones, case, digits
  static constexpr int NT = (NC + 15) / 16;
  static constexpr int NCP = 16 * NT;       // columns of the table
};
// Samples of one split at most (a multiple of 256): |(q - 128) d| <= 128 * 64 = 8192, and
// 8192 * 261,888 < 2^31.
constexpr size_t kSumsMaxRange = 261888;

struct SumsMExp {
  int e[2 * kMaxP + 1];  // E_f of each linear feature
};

typedef int i32x4s __attribute__((ext_vector_type(4)));

template <int P, int RT, int NW>
__global__ __launch_bounds__(64 * NW) void score_sums_mfma_kernel(
    LoadU8 load, const int8_t* __restrict__ limbs, const double* __restrict__ mu2p, SumsMExp fe,
    size_t npad, size_t nvar, size_t sps, double* __restrict__ part, size_t nvar_pad) {
  constexpr int NT = SumsM<P>::NT, NL = SumsM<P>::NL, NC = SumsM<P>::NC, NCP = SumsM<P>::NCP;
  constexpr int ND = SumsM<P>::NDIG, NS = Layout<P>::NS;
  // The NW waves share the sample range: each 256-sample block's table columns and mu2
  // are staged once per work-group in LDS (double-buffered, the next block's loaded into
  // registers during the current block's products).  Column stride 272 bytes: the 16
  // columns of a tile read by one instruction fall in distinct banks.
  constexpr int BSTR = 272, UNITS = NCP * 16, UPT = (UNITS + 64 * NW - 1) / (64 * NW);
  __shared__ __attribute__((aligned(16))) int8_t sb[2][NCP * BSTR];
  __shared__ __attribute__((aligned(16))) double sm[2][256];
  const int tid = threadIdx.x, lane = tid & 63, r = lane & 15, qd = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const size_t vb = ((size_t)blockIdx.x * NW + wave) * (16 * RT);
  const bool live = vb < nvar;  // wave-uniform; every wave takes part in the staging
  const size_t i0 = (size_t)blockIdx.y * sps, i1 = min(i0 + sps, npad);
  const uint8_t* rowp[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) rowp[t] = load.q + min(vb + 16 * t + r, nvar - 1) * load.ld;
  i32x4s acc[RT][NT], on[NT];
  double q2[RT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    on[nt] = i32x4s{0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < RT; ++t) acc[t][nt] = i32x4s{0, 0, 0, 0};
  }
#pragma unroll
  for (int t = 0; t < RT; ++t) q2[t] = 0.0;
  const i32x4s ones = {0x01010101, 0x01010101, 0x01010101, 0x01010101};
  // staging loads of block i: table units (16 bytes) u = tid + 256 k, one mu2 value
  constexpr int NTH = 64 * NW;
  static_assert(NTH >= 256, "a thread per staged mu2");
  i32x4s su[UPT];
  double smv = 0.0;
  auto fetch_stage = [&](size_t i) {
    const int8_t* src = limbs + (i / 256) * (size_t)(NCP * 256);
#pragma unroll
    for (int k = 0; k < UPT; ++k)
      if (UNITS % NTH == 0 || tid + NTH * k < UNITS)
        su[k] = *reinterpret_cast<const i32x4s*>(src + 16 * (tid + NTH * k));
    if (NTH == 256 || tid < 256) smv = mu2p[i + tid];
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
      const int u = tid + NTH * k;
      if (UNITS % NTH == 0 || u < UNITS)
        *reinterpret_cast<i32x4s*>(&sb[buf][(u >> 4) * BSTR + 16 * (u & 15)]) = su[k];
    }
    if (NTH == 256 || tid < 256) sm[buf][tid] = smv;
  };
  // dosage bytes of block i; a 16-byte unit past the row's ld is read from the row's start
  // instead (its samples are past n, where the table and mu2 are zero)
  auto fetch_rows = [&](size_t i, u32x4w (&dst)[RT][4]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t cs = i + 64 * j + 16 * qd;
        dst[t][j] = *reinterpret_cast<const u32x4w*>(rowp[t] + (cs < load.ld ? cs : 0));
      }
  };
  auto compute = [&](int buf, const u32x4w (&w)[RT][4]) {
    const int8_t* lb = &sb[buf][r * BSTR + 16 * qd];
    const double* mp = &sm[buf][16 * qd];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      i32x4s b[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        b[nt] = *reinterpret_cast<const i32x4s*>(lb + nt * 16 * BSTR + 64 * j);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        on[nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ones, b[nt], on[nt], 0, 0, 0);
      uint32_t nz = 0u;
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        const u32x4w x = w[t][j];
        nz |= x[0] | x[1] | x[2] | x[3];
        const i32x4s a = {(int)(x[0] ^ 0x80808080u), (int)(x[1] ^ 0x80808080u),
                          (int)(x[2] ^ 0x80808080u), (int)(x[3] ^ 0x80808080u)};
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[t][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[nt], acc[t][nt], 0, 0, 0);
      }
      if (__ballot(nz != 0u) != 0) {  // wave-uniform: skipped with no dosage in the step
        double m[16];
#pragma unroll
        for (int k = 0; k < 16; k += 2) {
          const f64x2 v2 = *reinterpret_cast<const f64x2*>(mp + 64 * j + k);
          m[k] = v2.x;
          m[k + 1] = v2.y;
        }
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            const uint32_t qb = (w[t][j][k >> 2] >> (8 * (k & 3))) & 0xffu;
            q2[t] = fma((double)(qb * qb), m[k], q2[t]);
          }
      }
    }
  };
  // Two blocks per trip, the row registers alternating; the next block's loads (the last
  // block reloads itself) are in flight during the current block's products, and its stage
  // is written to the other LDS buffer after them.
  u32x4w wa[RT][4], wb[RT][4];
  auto block = [&](size_t i, int buf, const u32x4w (&cur)[RT][4], u32x4w (&nxt)[RT][4]) {
    __syncthreads();  // block i staged; the other buffer no longer read
    const size_t in = i + 256 < i1 ? i + 256 : i;
    fetch_stage(in);
    fetch_rows(in, nxt);
    if (live) compute(buf, cur);
    store_stage(buf ^ 1);
  };
  if (i0 < i1) {
    fetch_stage(i0);
    fetch_rows(i0, wa);
    store_stage(0);
    for (size_t i = i0; i < i1; i += 512) {
      block(i, 0, wa, wb);
      if (i + 256 < i1) block(i + 256, 1, wb, wa);
    }
  }
  if (!live) return;
  const size_t split = blockIdx.y;
  const double sc1 = load.scale, sc2 = load.scale * load.scale;
  // q^2 mu2: the four lanes (r, qd) hold quarters of variant vb + 16 t + r
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    double v = q2[t];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    const size_t var = vb + 16 * t + lane;
    if (lane < 16 && var < nvar) part[(split * NS + 3 + P) * nvar_pad + var] = v * sc2;
  }
  // C layout: lane (c, g) holds rows 4 g .. 4 g + 3 of column c of each tile; a column's
  // scale 2^(E_f - 7 l), its feature, and the ones-row sum (equal in the four rows)
  const int c = lane & 15, g = lane >> 4;
  double csc[NT];
  int cf[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int col = 16 * nt + c;
    cf[nt] = (col >= 2 && col < NC) ? (col - 2) / ND : (col < 2 ? -1 - col : -3);
    csc[nt] = cf[nt] >= 0 ? ldexp(1.0, fe.e[cf[nt]] - 7 * ((col - 2) % ND + 1)) : 1.0;
  }
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double val[NL], cnt = 0.0, cse = 0.0;
#pragma unroll
      for (int f = 0; f < NL; ++f) val[f] = 0.0;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const double x = (double)acc[t][nt][i] + 128.0 * (double)on[nt][0];
        if (cf[nt] == -1) cnt = x;
        if (cf[nt] == -2) cse = x;
#pragma unroll
        for (int f = 0; f < NL; ++f) val[f] += cf[nt] == f ? x * csc[nt] : 0.0;
      }
#pragma unroll
      for (int m2 = 1; m2 < 16; m2 <<= 1) {
        cnt += __shfl_xor(cnt, m2);
        cse += __shfl_xor(cse, m2);
#pragma unroll
        for (int f = 0; f < NL; ++f) val[f] += __shfl_xor(val[f], m2);
      }
      const size_t var = vb + 16 * t + 4 * g + i;
      if (c == 0 && var < nvar) {
        double* o = part + split * NS * nvar_pad + var;
        o[0] = cnt * sc1;
        o[nvar_pad] = 0.0;
#pragma unroll
        for (int f = 0; f <= P; ++f) o[(2 + f) * nvar_pad] = val[f] * sc1;
#pragma unroll
        for (int j = 0; j < P; ++j) o[(4 + P + j) * nvar_pad] = val[P + 1 + j] * sc1;
        o[(4 + 2 * P) * nvar_pad] = cse * sc1;
        o[(5 + 2 * P) * nvar_pad] = 0.0;
      }
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
  if (load.vec) {
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
    s[5 + 2 * P] = ec.n_cases;
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
#pragma unroll
  for (int j = 0; j < P; ++j) {
    gres -= ag[j] * ec.xtr[j];
    q2 -= 2.0 * ag[j] * s[4 + P + j];
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
  // features of the learned root proposal; the carrier moments (3-5) are filled in by the
  // SPA setup for the variants that go to SPA
  r.spa_features[0] = (float)(vts > 0.0 ? score / sqrt(vts) : 0.0);
  r.spa_features[1] = (float)log(fmax(vt, 1e-300));
  r.spa_features[2] = (float)log1p(fmax(r.mac, 0.0));
  r.spa_features[3] = r.spa_features[4] = r.spa_features[5] = 0.0f;
  r.variance_ratio = vr;
  r.beta = vts > 1e-30 ? score / vts : 0.0;
  r.se_beta = vts > 1e-30 ? 1.0 / sqrt(vts) : __longlong_as_double(0x7ff0000000000000ll);
  if (ec.has_y) {
    const double sc = s[4 + 2 * P], nc = s[5 + 2 * P], sn = sum - sc;
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
    // q = S / sqrt(VR) + m1 and qinv = 2 m1 - q are completed by the SPA setup, which
    // sums m1 = sum mu g~ over the samples
    uint32_t slot = atomicAdd(n_gate, 1u);
    GateInfo gi;
    gi.var = (uint32_t)v;
    gi.pad = 0;
    gi.q = score / sqrt(vr);
    gi.qinv = 0.0;
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

// A search begun at a learned proposal (tinit != 0) that failed, or that ended at a root
// beyond kProposalMaxRoot, is run again from 0, once.
__device__ __forceinline__ bool proposal_abandoned(const RootState& s, double tinit) {
  return s.phase == 3 && tinit != 0.0 && s.spent == 0 &&
         (!s.converged || !(fabs(s.root) <= kProposalMaxRoot));
}

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
  // a learned proposal that fails or ends beyond kProposalMaxRoot is restarted once from 0,
  // the plain method's start, so that the outcome is the plain method's.
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
      if (proposal_abandoned(st[r], tinit[r])) {
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
// machine (root_after_eval, restart from 0 after an abandoned proposal).  A batch's few
// hundred gated variants then fill the GPU instead of occupying one work-group each.
// ---------------------------------------------------------------------------------
struct SpaVar {
  RootState st[2];
  double gpos, gneg, na_mu, na_sigma;
  double tinit[2];
  int done;          // both searches finished
  int bmin, bmax;    // smallest and largest dosage byte of the variant (byte-table passes)
  int split;         // byte-table passes over the carriers only (SpaZeroTab)
};
constexpr int kSpaPassBlock = 256;
constexpr int kSpaPart = 11;  // partials per (variant, slice): 8 sums, byte range, carriers

// Sums over every sample of an intercept-only model, as functions of s, for the byte-table
// passes of variants with few carriers: with x_i = rho_i e^s (rho_i = (1 - mu_i) / mu_i),
//   A(s) = sum 1 / (1 + x_i),  B(s) = sum x_i / (1 + x_i)^2,  D(s) = sum ln(1 - mu_i + mu_i e^-s).
// The samples with dosage byte 0 share g~_0, so their K'(t), K''(t) and K(zeta) terms are
// g~_0 A, g~_0^2 B and D at s = -g~_0 t, less the same terms of the carriers (bytes > 0),
// which the passes compute sample by sample with their own terms.  A, B and D / s are
// analytic in |Im s| < pi (their singularities sit at s = -ln rho_i +- i pi; D(0) = 0); on
// [s_lo, s_hi] each unit interval holds Chebyshev series of kZtN terms fitted at its
// Chebyshev points to the exact sums (series error below 1e-18 relative for a Bernstein
// ellipse of semi-minor axis 2, so the fit carries the rounding of the node sums).  D is
// kept as D / s, so that near s = 0, where D ~ -s sum mu and K(zeta) - zeta q is small,
// its error scales with s.  Outside, where every x_i is below e^-40 or above e^40, the first
// terms of the expansions in x_i (or 1 / x_i) are exact to double precision.
constexpr int kZtN = 21;
constexpr double kZtMargin = 40.0;
struct SpaZeroTab {
  const double* coef;  // [piece][function A, B, D][kZtN]
  double s_lo, s_hi;
  int npiece, ok;
  double n, rho_sum, rho_inv_sum, lnmu_sum, ln1mmu_sum;
};
__device__ __forceinline__ double zt_eval(const SpaZeroTab& T, double s, int f) {
  if (s < T.s_lo) {
    const double x = exp(s);
    if (f == 0) return T.n - x * T.rho_sum;
    if (f == 1) return x * T.rho_sum;
    return fma(x, T.rho_sum, T.lnmu_sum - T.n * s);
  }
  if (!(s < T.s_hi)) {
    const double x = exp(-s);
    return f == 2 ? fma(x, T.rho_inv_sum, T.ln1mmu_sum) : x * T.rho_inv_sum;
  }
  const double u = s - T.s_lo;
  int p = (int)u;
  p = p >= T.npiece ? T.npiece - 1 : p;
  const double y = 2.0 * (u - (double)p) - 1.0;
  const double* c = T.coef + ((size_t)p * 3 + f) * kZtN;
  double b1 = 0.0, b2 = 0.0;
  for (int j = kZtN - 1; j >= 1; --j) {
    const double b0 = fma(2.0 * y, b1, c[j] - b2);
    b2 = b1;
    b1 = b0;
  }
  const double v = fma(y, b1, 0.5 * c[0] - b2);
  return f == 2 ? v * s : v;
}
// Exact A, B and D / s at the Chebyshev points of the table (one work-group per point;
// ln(1 - mu + mu e^-s) as log1p(mu expm1(-s)), accurate near s = 0).
__global__ __launch_bounds__(256) void spa_zero_nodes_kernel(const double* __restrict__ odds,
                                                             const double* __restrict__ mu,
                                                             size_t n,
                                                             const double* __restrict__ nodes,
                                                             double* __restrict__ out) {
  const double s = nodes[blockIdx.x], es = exp(s), em1 = expm1(-s);
  double v[3] = {0.0, 0.0, 0.0};
  for (size_t i = threadIdx.x; i < n; i += 256) {
    const double x = odds[i] * es;
    const double u = 1.0 / (1.0 + x);
    v[0] += u;
    v[1] = fma(x * u, u, v[1]);
    v[2] += log1p(mu[i] * em1);
  }
  __shared__ double sh[3][256];
  for (int f = 0; f < 3; ++f) sh[f][threadIdx.x] = v[f];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int f = 0; f < 3; ++f) sh[f][threadIdx.x] += sh[f][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3)
    out[3 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0] / (threadIdx.x == 2 ? s : 1.0);
}
// Carriers queued per wave in the split passes (half a step queues at most 128 behind at
// most 63).
constexpr uint32_t kCarrierQ = 256;
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int kSpaMaxPasses = 8 * 1000 + 16;  // the guard of spa_kernel's pass loop

// Sum of v[0..K) over the NT threads of the block, returned in every thread: a butterfly
// within each wave, then the waves' totals in wave order (a fixed order).  sh holds K * NT
// / 64 doubles.
template <int NT, int K>
__device__ __forceinline__ void block_sum_t(double* v, double* sh) {
  constexpr int NWV = NT / 64;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < K; ++k) sh[k * NWV + wv] = v[k];
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    double t = sh[k * NWV];
    for (int w = 1; w < NWV; ++w) t += sh[k * NWV + w];
    v[k] = t;
  }
  __syncthreads();
}

// 1/d for finite d > 0: the hardware estimate refined by two Newton steps (d >= mu >= 1e-10
// here; over 2^28 sampled d = mu + (1 - mu) e^x it equals the IEEE quotient in every case);
// recip_nr also gives 1/inf = 0 as IEEE division does.
__device__ __forceinline__ double recip_nr_finite(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double recip_nr(double d) {
  const double r = recip_nr_finite(d);
  return isinf(d) ? 0.0 : r;
}

// Running product for a sum of logs: sum_i ln x_i = ln(P) + E ln 2 with P a product of
// mantissas and E an integer.  mul() splits each factor x > 0 by frexp (exact) into a
// mantissa in [0.5, 1) and an exponent, norm() brings P back to [0.5, 1) and is called at
// least once every 8 factors, so P stays in [2^-9, 1).  One rounding of relative size 2^-53
// per factor replaces a logarithm per factor: at n = 400,000 the sum moves by at most
// n 2^-53 ~ 4e-11 (about 1e-13 for random rounding).  An infinite or NaN factor makes P
// infinite or NaN, as the sum of logs would be.
struct LogProd {
  double p = 1.0;
  int e = 0;
  __device__ __forceinline__ void mul(double x) {
    int ex;
    p *= frexp(x, &ex);
    e += ex;
  }
  __device__ __forceinline__ void norm() {
    int ex;
    p = frexp(p, &ex);
    e += ex;
  }
  __device__ __forceinline__ double log_value() const {
    return log(p) + (double)e * 0.69314718055994530942;  // ln 2
  }
};

// The Newton passes (KIND 1) work on the list of variants still searching: the step kernel
// of each pass appends the variants that go on to the other half of a double list, and a
// pass splits each listed variant into S slices with S chosen from the list's length, so
// that a pass has about kSpaPassTarget work-groups however many variants remain (a pass
// with few variants left takes as long as its share of the work, not as long as one slice
// of a full batch).  A variant-slice's partials sit at its position in the pass; the
// variant's sum over its slices is taken in slice order, so a variant's result depends only
// on the list's length, never on its place in the list.
constexpr uint32_t kSpaPassTarget = 4096;
struct SpaList {
  uint32_t* cnt;   // [2] variants in the list of pass parity 0 / 1
  uint32_t* list;  // [2][ng] gate slots
  uint32_t ng;
  int pass;
};
__host__ __device__ __forceinline__ void spa_pass_dims(uint32_t active, size_t n, uint32_t* S,
                                                       size_t* slice) {
  size_t s = active == 0 ? 1 : kSpaPassTarget / active;
  s = s < 1 ? 1 : (s > 64 ? 64 : s);
  const size_t smax = (n + 2047) / 2048;  // slices of at least 2,048 samples
  s = s > smax ? smax : s;
  if (s < 1) s = 1;
  const size_t sl = ((n + s - 1) / s + 255) / 256 * 256;
  *slice = sl;
  *S = (uint32_t)((n + sl - 1) / sl);
}
// What a pass work-group computes: the gate slot, its sample range and the position of its
// partials.  KIND 0 and 2: grid (gated variants, S slices of `slice`); KIND 1: the list.
struct SpaWork {
  uint32_t slot;
  size_t i0, i1, pos;
  bool first;  // the variant's first slice
  bool live;
};
template <int KIND>
__device__ __forceinline__ SpaWork spa_work(const uint32_t* __restrict__ n_gate, size_t n,
                                            size_t slice, const SpaList& sl) {
  SpaWork w;
  if constexpr (KIND != 1) {
    w.slot = blockIdx.x;
    w.live = w.slot < *n_gate;
    w.i0 = (size_t)blockIdx.y * slice;
    w.pos = (size_t)blockIdx.x * gridDim.y + blockIdx.y;
    w.first = blockIdx.y == 0;
  } else {
    const int h = sl.pass & 1;
    const uint32_t active = sl.cnt[h];
    uint32_t S;
    size_t slc;
    spa_pass_dims(active, n, &S, &slc);
    const uint32_t b = blockIdx.x;
    w.live = b < active * S;
    const uint32_t j = w.live ? b / S : 0u, q = w.live ? b % S : 0u;
    w.slot = w.live ? sl.list[(size_t)h * sl.ng + j] : 0u;
    w.i0 = (size_t)q * slc;
    slice = slc;
    w.pos = b;
    w.first = q == 0;
  }
  w.i1 = min(w.i0 + slice, n);
  return w;
}

// KIND 0: g_pos, g_neg, the NA moments, m1 = sum mu g~ (single_variant.rs:189-192) and the
// carrier moments sum mu, sum mu(1-mu), sum mu^2(1-mu) over samples with dosage > 0 (the
// learned proposal's inputs); 1: K'(t) and K''(t) sums at both pending points; 2: K(zeta)
// and K''(zeta) sums at both roots (variants whose two searches converged).
template <class G, int KIND>
__global__ __launch_bounds__(kSpaPassBlock) void spa_pass_kernel(
    G gt, const double* __restrict__ mu, size_t n, const GateInfo* __restrict__ gate,
    const uint32_t* __restrict__ n_gate, int fast, size_t slice,
    const SpaVar* __restrict__ vars, double* __restrict__ part, SpaList sl) {
  constexpr int K = KIND == 0 ? 8 : 4;
  if (KIND == 1 && blockIdx.x == 0 && threadIdx.x == 0) sl.cnt[(sl.pass + 1) & 1] = 0u;
  const SpaWork wk = spa_work<KIND>(n_gate, n, slice, sl);
  if (!wk.live) return;
  const uint32_t slot = wk.slot;
  if (KIND == 2 && !(vars[slot].st[0].converged && vars[slot].st[1].converged)) return;
  __shared__ double sh[K * kSpaPassBlock / 64];
  __shared__ GateInfo g;
  if (threadIdx.x == 0) g = gate[slot];
  __syncthreads();
  const size_t v = g.var;
  const double thr = fast ? 1e-10 : -1.0;  // fast: |g~| > 1e-10 enters K (fast.rs:191-200)
  double t0 = 0.0, t1 = 0.0;
  const bool act0 = KIND != 1 || vars[slot].st[0].phase != 3;
  const bool act1 = KIND != 1 || vars[slot].st[1].phase != 3;
  if (KIND == 1) {
    t0 = vars[slot].st[0].pend;
    t1 = vars[slot].st[1].pend;
  } else if (KIND == 2) {
    t0 = vars[slot].st[0].root;
    t1 = vars[slot].st[1].root;
  }
  const size_t i0 = wk.i0, i1 = wk.i1;
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  LogProd lk0, lk1;  // KIND 2: K at the roots as the logs of running products
  for (size_t i = i0 + threadIdx.x; i < i1; i += kSpaPassBlock) {
    const double m = mu[i];
    if (KIND == 0) {
      const bool carrier = gt.load(v, i) > 0.0;
      const double gi = gt(v, i, g.ag);
      if (gi > 0.0) a[0] += gi;
      if (gi < 0.0) a[1] += gi;
      const double m2 = m * (1.0 - m);
      if (!(fabs(gi) > thr)) {
        a[2] += m * gi;
        a[3] += m2 * gi * gi;
      }
      a[4] = fma(m, gi, a[4]);
      if (carrier) {
        a[5] += m;
        a[6] += m2;
        a[7] = fma(m, m2, a[7]);
      }
      continue;
    }
    const double gi = gt(v, i, g.ag);
    if (!(fabs(gi) > thr)) continue;
    const double om = 1.0 - m;
    const double mg = m * gi, w2 = om * mg * gi;
    if (KIND == 1) {  // a finished root's terms are skipped (uniform over the block)
      if (act0) {
        const double e0 = exp(-gi * t0), r0 = recip_nr(fma(om, e0, m));
        a[0] = fma(mg, r0, a[0]);
        a[1] = fma(w2 * e0, r0 * r0, a[1]);
      }
      if (act1) {
        const double e1 = exp(-gi * t1), r1 = recip_nr(fma(om, e1, m));
        a[2] = fma(mg, r1, a[2]);
        a[3] = fma(w2 * e1, r1 * r1, a[3]);
      }
      continue;
    }
    const double e0 = exp(-gi * t0), e1 = exp(-gi * t1);
    const double r0 = recip_nr(fma(om, e0, m)), r1 = recip_nr(fma(om, e1, m));
    {
      // K at the roots is the sum of ln(1 - m + m e^{g zeta}) (binary.rs:148-151)
      lk0.mul(fma(m, exp(gi * t0), om));
      lk1.mul(fma(m, exp(gi * t1), om));
      lk0.norm();
      lk1.norm();
    }
    a[1] = fma(w2 * e0, r0 * r0, a[1]);
    a[3] = fma(w2 * e1, r1 * r1, a[3]);
  }
  if (KIND == 2) {
    a[0] = lk0.log_value();
    a[2] = lk1.log_value();
  }
  block_sum_t<kSpaPassBlock, K>(a, sh);
  if (threadIdx.x == 0) {
    double* o = part + wk.pos * kSpaPart;
    for (int k = 0; k < K; ++k) o[k] = a[k];
    if (KIND == 0) {
      o[8] = 0.0;
      o[9] = 255.0;
      o[10] = 0.0;
    }
  }
}

// The passes of spa_pass_kernel for uint8 dosages of an intercept-only model, where g~
// depends only on the dosage byte: per byte, LDS holds g~ and the exponentials of the pass
// (g~ = 0 and unit exponentials for bytes whose samples do not enter K under the fast
// rule), and each thread takes four consecutive samples per step (one 32-bit dosage load and
// two 16-byte mu loads), with the next step's loads in flight.
template <int KIND>
__global__ __launch_bounds__(kSpaPassBlock) void spa_pass_tab_kernel(
    LoadU8 ld8, const double* __restrict__ mu, size_t n, const GateInfo* __restrict__ gate,
    const uint32_t* __restrict__ n_gate, int fast, size_t slice,
    const SpaVar* __restrict__ vars, double* __restrict__ part, SpaList sl, SpaZeroTab zt) {
  constexpr int K = KIND == 0 ? 8 : 4;
  constexpr int TW = KIND == 0 ? 2 : (KIND == 1 ? 3 : 6);  // doubles per byte entry
  if (KIND == 1 && blockIdx.x == 0 && threadIdx.x == 0) sl.cnt[(sl.pass + 1) & 1] = 0u;
  const SpaWork wk = spa_work<KIND>(n_gate, n, slice, sl);
  if (!wk.live) return;
  const uint32_t slot = wk.slot;
  if (KIND == 2 && !(vars[slot].st[0].converged && vars[slot].st[1].converged)) return;
  __shared__ double sh[K * kSpaPassBlock / 64];
  __shared__ __attribute__((aligned(16))) double tab[256 * TW];
  __shared__ GateInfo g;
  if (threadIdx.x == 0) g = gate[slot];
  __syncthreads();
  const double thr = fast ? 1e-10 : -1.0;  // fast: |g~| > 1e-10 enters K (fast.rs:191-200)
  {
    static_assert(kSpaPassBlock == 256, "one table entry per thread");
    const int b = threadIdx.x;
    const double gb = (double)b * ld8.scale - g.ag[0];
    const bool incl = fabs(gb) > thr;
    double* e = tab + b * TW;
    if (KIND == 0) {
      e[0] = gb;
      e[1] = incl ? 1.0 : 0.0;
    } else {
      const double t0 = KIND == 1 ? vars[slot].st[0].pend : vars[slot].st[0].root;
      const double t1 = KIND == 1 ? vars[slot].st[1].pend : vars[slot].st[1].root;
      e[0] = incl ? gb : 0.0;
      e[1] = incl ? exp(-gb * t0) : 1.0;
      e[2] = incl ? exp(-gb * t1) : 1.0;
      if (KIND == 1) {
        // an overflowing e^{-g t} makes that sample's K' term 0 and its K'' term NaN
        // (m g / inf, inf / inf): the capped value (with odds at most 1e12, 1 + odds e stays
        // finite) gives K' and K'' terms below 1e-283 relative to g, and the NaN is added
        // once per variant below
        e[1] = fmin(e[1], 1e295);
        e[2] = fmin(e[2], 1e295);
      }
      if (KIND == 2) {
        e[3] = incl ? exp(gb * t0) : 1.0;
        e[4] = incl ? exp(gb * t1) : 1.0;
        e[5] = incl ? 1.0 : 0.0;
      }
    }
  }
  __syncthreads();
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  LogProd lk0, lk1;  // KIND 2: K at the roots as the logs of running products
  uint32_t bmin = 255u, bmax = 0u, ncar = 0u;
  // act: bit r set while root r still searches (KIND 1 skips the finished root's terms)
  auto body = [&](uint32_t by, double m, auto act) {
    constexpr int ACT = decltype(act)::value;
    const double* e = tab + by * TW;
    const double om = 1.0 - m;
    if (KIND == 0) {
      const double gi = e[0];
      if (gi > 0.0) a[0] += gi;
      if (gi < 0.0) a[1] += gi;
      const double m2 = m * om;
      if (e[1] == 0.0) {
        a[2] += m * gi;
        a[3] += m2 * gi * gi;
      }
      a[4] = fma(m, gi, a[4]);
      if (by != 0u) {
        a[5] += m;
        a[6] += m2;
        a[7] = fma(m, m2, a[7]);
      }
      bmin = min(bmin, by);
      ncar += by != 0u ? 1u : 0u;
      bmax = max(bmax, by);
    } else if (KIND == 1) {
      // here m holds the odds rho = (1 - mu) / mu: with u = mu / ((1 - mu) e + mu) =
      // 1 / (1 + rho e), the K' term mu g / ((1 - mu) e + mu) is g u and the K'' term
      // (1 - mu) mu g^2 e / ((1 - mu) e + mu)^2 is g u (g - g u)
      const double gk = e[0];
      if constexpr ((ACT & 1) != 0) {
        const double g0 = gk * recip_nr_finite(fma(m, e[1], 1.0));
        a[0] += g0;
        a[1] = fma(g0, gk - g0, a[1]);
      }
      if constexpr ((ACT & 2) != 0) {
        const double g1 = gk * recip_nr_finite(fma(m, e[2], 1.0));
        a[2] += g1;
        a[3] = fma(g1, gk - g1, a[3]);
      }
    } else {
      const double gk = e[0], e0 = e[1], e1 = e[2];
      const double mg = m * gk, w2 = om * mg * gk;
      const double r0 = recip_nr(fma(om, e0, m));
      const double r1 = recip_nr(fma(om, e1, m));
      {
        // K at the roots is the sum of ln(1 - m + m e^{g zeta}) over the included bytes
        // (binary.rs:148-151); an excluded byte's factor is 1
        lk0.mul(e[5] != 0.0 ? fma(m, e[3], om) : 1.0);
        lk1.mul(e[5] != 0.0 ? fma(m, e[4], om) : 1.0);
      }
      a[1] = fma(w2 * e0, r0 * r0, a[1]);
      a[3] = fma(w2 * e1, r1 * r1, a[3]);
    }
  };
  const uint8_t* row = ld8.q + (size_t)g.var * ld8.ld;
  const size_t i0 = wk.i0, i1 = wk.i1;
  const size_t i4 = i0 + ((i1 - i0) & ~size_t(3));
  constexpr size_t kStep = 4 * kSpaPassBlock;
  auto sweep = [&](auto act) {
    size_t tail = i0;
    if (ld8.vec) {  // 16-byte aligned rows; i0 is a multiple of 256
      size_t i = i0 + 4 * threadIdx.x;
      uint32_t wn = 0u;
      f64x2 ma = {0.0, 0.0}, mb = {0.0, 0.0};
      if (i < i4) {
        wn = *reinterpret_cast<const uint32_t*>(row + i);
        ma = *reinterpret_cast<const f64x2*>(mu + i);
        mb = *reinterpret_cast<const f64x2*>(mu + i + 2);
      }
      for (; i < i4; i += kStep) {
        const uint32_t w = wn;
        const f64x2 ca = ma, cb = mb;
        if (i + kStep < i4) {
          wn = *reinterpret_cast<const uint32_t*>(row + i + kStep);
          ma = *reinterpret_cast<const f64x2*>(mu + i + kStep);
          mb = *reinterpret_cast<const f64x2*>(mu + i + kStep + 2);
        }
        body(w & 0xffu, ca.x, act);
        body((w >> 8) & 0xffu, ca.y, act);
        body((w >> 16) & 0xffu, cb.x, act);
        body(w >> 24, cb.y, act);
        if (KIND == 2) {
          lk0.norm();
          lk1.norm();
        }
      }
      tail = i4;
    }
    for (size_t j = tail + threadIdx.x; j < i1; j += kSpaPassBlock) {
      body(row[j], mu[j], act);
      if (KIND == 2) {
        lk0.norm();
        lk1.norm();
      }
    }
  };
  const bool act0 = KIND != 1 || vars[slot].st[0].phase != 3;
  const bool act1 = KIND != 1 || vars[slot].st[1].phase != 3;
  bool split = false;
  if constexpr (KIND != 0) split = vars[slot].split != 0;  // uniform over the work-group
  if (!split) {
    if (act0 && act1) sweep(std::integral_constant<int, 3>{});
    else if (act0) sweep(std::integral_constant<int, 1>{});
    else sweep(std::integral_constant<int, 2>{});
    if (KIND == 2) {
      a[0] = lk0.log_value();
      a[2] = lk1.log_value();
    }
  } else if constexpr (KIND != 0) {
    // Carriers only (dosage byte > 0): each wave queues its carriers' bytes and mu (odds)
    // in LDS and takes them 64 at a time, one per lane, adding the carrier's own terms and
    // subtracting the terms it would have with byte 0's g~_0; byte 0's terms over all
    // samples come from the table (SpaZeroTab), once per variant.
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const double* ez = tab;  // byte 0's entry
    const bool zin = KIND == 1 ? ez[0] != 0.0 : ez[5] != 0.0;
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    LogProd lz0, lz1;
    auto carrier = [&](uint32_t by, double m, auto act) {
      constexpr int ACT = decltype(act)::value;
      body(by, m, act);
      if (KIND == 2) {
        lk0.norm();
        lk1.norm();
      }
      if (!zin) return;
      const double gz = ez[0];
      if constexpr (KIND == 1) {
        if constexpr ((ACT & 1) != 0) {
          const double z0 = gz * recip_nr_finite(fma(m, ez[1], 1.0));
          c[0] += z0;
          c[1] = fma(z0, gz - z0, c[1]);
        }
        if constexpr ((ACT & 2) != 0) {
          const double z1 = gz * recip_nr_finite(fma(m, ez[2], 1.0));
          c[2] += z1;
          c[3] = fma(z1, gz - z1, c[3]);
        }
      } else {
        const double om = 1.0 - m, mg = m * gz, w2 = om * mg * gz;
        const double r0 = recip_nr(fma(om, ez[1], m)), r1 = recip_nr(fma(om, ez[2], m));
        lz0.mul(fma(m, ez[3], om));
        lz1.mul(fma(m, ez[4], om));
        lz0.norm();
        lz1.norm();
        c[1] = fma(w2 * ez[1], r0 * r0, c[1]);
        c[3] = fma(w2 * ez[2], r1 * r1, c[3]);
      }
    };
    // carrier queue of each wave: the sample's mu (odds in the Newton passes) and byte
    __shared__ double cqm[kSpaPassBlock / 64][kCarrierQ];
    __shared__ uint8_t cqb[kSpaPassBlock / 64][kCarrierQ];
    auto split_sweep = [&](auto act) {
      uint32_t head = 0u, tail = 0u;  // wave-uniform
      auto drain = [&](uint32_t cnt) {
        wave_sync();
        if ((uint32_t)ln < cnt) {
          const uint32_t e = (head + (uint32_t)ln) & (kCarrierQ - 1);
          carrier(cqb[wv][e], cqm[wv][e], act);
        }
        head += cnt;
        wave_sync();
      };
      auto enqueue = [&](uint32_t w, const double* m4) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          uint32_t t = tail;
#pragma unroll
          for (int k = 2 * h; k < 2 * h + 2; ++k) {
            const uint32_t by = (w >> (8 * k)) & 0xffu;
            const uint64_t mk = __ballot(by != 0u);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
            if (by != 0u) {
              const uint32_t e = (t + below) & (kCarrierQ - 1);
              cqm[wv][e] = m4[k];
              cqb[wv][e] = (uint8_t)by;
            }
            t += (uint32_t)__builtin_popcountll(mk);
          }
          tail = t;
          while (tail - head >= 64u) drain(64u);
        }
      };
      // wave wv takes samples base .. base + 255 of every 1,024 (4 per lane); the next
      // step's bytes and mu are in flight while a step is queued and its carriers taken
      size_t base = i0 + 256 * (size_t)wv;
      if (ld8.vec) {
        const size_t i4 = i0 + ((i1 - i0) & ~size_t(3));
        auto ld = [&](size_t i, uint32_t& w, f64x2& ma, f64x2& mb) {
          if (i < i4) {
            w = *reinterpret_cast<const uint32_t*>(row + i);
            ma = *reinterpret_cast<const f64x2*>(mu + i);
            mb = *reinterpret_cast<const f64x2*>(mu + i + 2);
          } else {
            w = 0u;
          }
        };
        uint32_t wn = 0u;
        f64x2 ma = {0.0, 0.0}, mb = {0.0, 0.0};
        if (base < i1) ld(base + 4 * (size_t)ln, wn, ma, mb);
        for (; base < i1; base += kStep) {
          const uint32_t w = wn;
          const double m4[4] = {ma.x, ma.y, mb.x, mb.y};
          if (base + kStep < i1) ld(base + kStep + 4 * (size_t)ln, wn, ma, mb);
          enqueue(w, m4);
        }
        // samples past the last whole group of four, one per lane
        for (size_t j = i4 + (size_t)threadIdx.x; j < i1; j += kSpaPassBlock)
          if (row[j] != 0u) carrier(row[j], mu[j], act);
      } else {
        for (; base < i1; base += kStep) {
          const size_t i = base + 4 * (size_t)ln;
          uint32_t w = 0u;
          double m4[4] = {0.0, 0.0, 0.0, 0.0};
          for (int k = 0; k < 4; ++k)
            if (i + k < i1) {
              w |= (uint32_t)row[i + k] << (8 * k);
              m4[k] = mu[i + k];
            }
          enqueue(w, m4);
        }
      }
      drain(tail - head);
    };
    if (act0 && act1) split_sweep(std::integral_constant<int, 3>{});
    else if (act0) split_sweep(std::integral_constant<int, 1>{});
    else split_sweep(std::integral_constant<int, 2>{});
    const bool lead = wk.first && threadIdx.x == 0;
    const double gz = ez[0];
    if constexpr (KIND == 1) {
#pragma unroll
      for (int k = 0; k < 4; ++k) a[k] -= c[k];
      if (lead && zin) {
        if (act0) {
          const double s0 = -gz * vars[slot].st[0].pend;
          a[0] = fma(gz, zt_eval(zt, s0, 0), a[0]);
          a[1] = fma(gz * gz, zt_eval(zt, s0, 1), a[1]);
        }
        if (act1) {
          const double s1 = -gz * vars[slot].st[1].pend;
          a[2] = fma(gz, zt_eval(zt, s1, 0), a[2]);
          a[3] = fma(gz * gz, zt_eval(zt, s1, 1), a[3]);
        }
      }
    } else {
      a[0] = lk0.log_value() - lz0.log_value();
      a[2] = lk1.log_value() - lz1.log_value();
      a[1] -= c[1];
      a[3] -= c[3];
      if (zin) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const double zeta = vars[slot].st[r].root;
          // an infinite factor 1 - mu + mu e^{g zeta} of a byte-0 sample makes K infinite,
          // and an infinite e^{-g zeta} its K'' NaN (inf * 0), as in the per-sample sums
          if (isinf(exp(gz * zeta))) a[2 * r] = lead ? __longlong_as_double(0x7ff0000000000000ll) : 0.0;
          else if (lead) a[2 * r] += zt_eval(zt, -gz * zeta, 2);
          if (lead)
            a[2 * r + 1] += isinf(exp(-gz * zeta)) ? __longlong_as_double(0x7ff8000000000000ll)
                                                    : gz * gz * zt_eval(zt, -gz * zeta, 1);
        }
      }
    }
  }
  __shared__ uint32_t brange[3];
  if (KIND == 0) {
    if (threadIdx.x == 0) {
      brange[0] = 255u;
      brange[1] = 0u;
      brange[2] = 0u;
    }
    __syncthreads();
    atomicMin(&brange[0], bmin);
    atomicMax(&brange[1], bmax);
    atomicAdd(&brange[2], ncar);
  }
  block_sum_t<kSpaPassBlock, K>(a, sh);
  if (threadIdx.x == 0) {
    double* o = part + wk.pos * kSpaPart;
    if (KIND == 1 && wk.first) {
      // g~ is increasing in the byte, so e^{-g t} is largest at the smallest or the largest
      // byte of the variant: K'' is NaN when either overflows (an included byte)
      const SpaVar& V = vars[slot];
      for (int r = 0; r < 2; ++r) {
        const double t = V.st[r].pend;
        bool nan = false;
        for (int b : {V.bmin, V.bmax}) {
          const double gb = (double)b * ld8.scale - g.ag[0];
          if (fabs(gb) > thr && isinf(exp(-gb * t))) nan = true;
        }
        if (nan) a[2 * r + 1] += __longlong_as_double(0x7ff8000000000000ll);
      }
    }
    for (int k = 0; k < K; ++k) o[k] = a[k];
    if (KIND == 0) {
      o[8] = (double)brange[0];
      o[9] = (double)brange[1];
      o[10] = (double)brange[2];
    }
  }
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

// Totals of the setup pass: q = S / sqrt(VR) + m1, qinv = 2 m1 - q (single_variant.rs:
// 189-196), the support bounds, the NA moments and the carrier-moment features.
// Variants whose carriers are at most kSplitCarriers of the samples take the carrier-only
// byte-table passes (when the engine has the zero-byte table).
constexpr double kSplitCarriers = 0.35;
// One wave per variant: lane q holds slice q's partials (S <= 64), added in a fixed
// butterfly order.
__global__ __launch_bounds__(256) void spa_setup_step(GateInfo* __restrict__ gate,
                                                      const uint32_t* __restrict__ n_gate, int S,
                                                      const double* __restrict__ part,
                                                      SpaVar* __restrict__ vars,
                                                      saige_variant_result* __restrict__ out,
                                                      size_t n, int split_ok) {
  const uint32_t slot = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (slot >= *n_gate) return;  // uniform over the wave
  double t[kSpaPart];
  const bool have = lane < S;
  const double* pp = part + ((size_t)slot * S + (have ? lane : 0)) * kSpaPart;
#pragma unroll
  for (int k = 0; k < kSpaPart; ++k) t[k] = have ? pp[k] : 0.0;
  double bmn = have ? t[8] : 255.0, bmx = have ? t[9] : 0.0;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int k = 0; k < kSpaPart; ++k)
      if (k != 8 && k != 9) t[k] += __shfl_xor(t[k], m);
    bmn = fmin(bmn, __shfl_xor(bmn, m));
    bmx = fmax(bmx, __shfl_xor(bmx, m));
  }
  if (lane != 0) return;
  SpaVar V = vars[slot];
  V.split = (split_ok && t[10] <= kSplitCarriers * (double)n) ? 1 : 0;
  V.gpos = t[0];
  V.gneg = t[1];
  V.na_mu = t[2];
  V.na_sigma = t[3];
  V.tinit[0] = V.tinit[1] = 0.0;  // fast.rs:217-218
  V.bmin = (int)bmn;
  V.bmax = (int)bmx;
  vars[slot] = V;
  const double m1 = t[4];
  GateInfo& g = gate[slot];
  g.q = g.q + m1;
  g.qinv = 2.0 * m1 - g.q;
  for (int j = 0; j < 3; ++j) {
    const float f = (float)log(t[5 + j] + 1e-8);
    g.feat[3 + j] = (double)f;
    if (out) out[g.var].spa_features[3 + j] = f;
  }
}

// Begin both searches at 0 (fast.rs:217-218) or at the learned proposal; a q outside the
// support (fast.rs:67-73) has its root at infinity.
__global__ void spa_start_step(const GateInfo* __restrict__ gate,
                               const uint32_t* __restrict__ n_gate, SpaVar* __restrict__ vars,
                               SpaList sl) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= *n_gate) return;
  SpaVar V = vars[slot];
  const double q[2] = {gate[slot].q, gate[slot].qinv};
  for (int r = 0; r < 2; ++r) {
    root_start(V.st[r], V.tinit[r]);
    V.st[r].spent = 0;
    if (q[r] >= V.gpos || q[r] <= V.gneg) {
      V.st[r].root = __longlong_as_double(0x7ff0000000000000ll);
      V.st[r].converged = 1;
      V.st[r].phase = 3;
    }
  }
  V.done = (V.st[0].phase == 3 && V.st[1].phase == 3) ? 1 : 0;
  vars[slot] = V;
  if (!V.done) sl.list[atomicAdd(&sl.cnt[0], 1u)] = slot;  // the list of pass 0
}

// Step of Newton pass `pass` (sl.pass): one wave per listed variant. Lane q loads slice q's
// partials (S <= 64) and the wave adds them in a fixed butterfly order (lane 0's totals are
// used, so a variant's sums depend only on S); lane 0 then advances both searches and lists
// the variant for the next pass if it goes on.
__global__ __launch_bounds__(256) void spa_eval_step(const GateInfo* __restrict__ gate, size_t n,
                                                     const double* __restrict__ part, double tol,
                                                     SpaVar* __restrict__ vars, SpaList sl) {
  const uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int h = sl.pass & 1;
  const uint32_t active = sl.cnt[h];
  if (j >= active) return;  // uniform over the wave
  uint32_t S;
  size_t slc;
  spa_pass_dims(active, n, &S, &slc);
  double sum[4] = {0.0, 0.0, 0.0, 0.0};
  if ((uint32_t)lane < S) {
    const double* pp = part + ((size_t)j * S + lane) * kSpaPart;
#pragma unroll
    for (int k = 0; k < 4; ++k) sum[k] = pp[k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) sum[k] += __shfl_xor(sum[k], off);
  if (lane != 0) return;
  const uint32_t slot = sl.list[(size_t)h * sl.ng + j];
  SpaVar V = vars[slot];
  const double q[2] = {gate[slot].q, gate[slot].qinv};
  const int max_iter = 1000;
  for (int r = 0; r < 2; ++r) {
    if (V.st[r].phase == 3) continue;
    const double tt = V.st[r].pend;
    const double k1 = sum[2 * r] + V.na_mu + V.na_sigma * tt - q[r];
    const double k2 = sum[2 * r + 1] + V.na_sigma;
    root_after_eval(V.st[r], k1, k2, tol, max_iter);
    if (proposal_abandoned(V.st[r], V.tinit[r])) {
      V.st[r].spent = V.st[r].it;
      root_start(V.st[r], 0.0);
    }
  }
  V.done = (V.st[0].phase == 3 && V.st[1].phase == 3) ? 1 : 0;
  vars[slot] = V;
  if (!V.done) sl.list[(size_t)(h ^ 1) * sl.ng + atomicAdd(&sl.cnt[h ^ 1], 1u)] = slot;
}

// saddle-point probabilities (fast.rs:135-180 and :220-241) and the outputs
// One wave per variant: the slices' K and K'' sums added by a butterfly (S <= 64), then lane
// 0 finishes the variant.
__global__ __launch_bounds__(256) void spa_final_step(const GateInfo* __restrict__ gate,
                                                      const uint32_t* __restrict__ n_gate, int S,
                                                      const double* __restrict__ part,
                                                      const SpaVar* __restrict__ vars,
                                                      saige_variant_result* __restrict__ out) {
  const uint32_t slot = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (slot >= *n_gate) return;  // uniform over the wave
  double ks[4] = {0.0, 0.0, 0.0, 0.0};
  if (lane < S) {
    const double* pp = part + ((size_t)slot * S + lane) * kSpaPart;
#pragma unroll
    for (int k = 0; k < 4; ++k) ks[k] = pp[k];
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) ks[k] += __shfl_xor(ks[k], m);
  if (lane != 0) return;
  const SpaVar V = vars[slot];
  const GateInfo g = gate[slot];
  const double q[2] = {g.q, g.qinv};
  const bool both = V.st[0].converged && V.st[1].converged;
  double p[2] = {0.0, 0.0};
  bool ok[2] = {false, false};
  if (both) {
    for (int r = 0; r < 2; ++r) {
      const double zeta = V.st[r].root;
      const double k0 = ks[2 * r] + V.na_mu * zeta + 0.5 * V.na_sigma * zeta * zeta;
      const double k2 = ks[2 * r + 1] + V.na_sigma;
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
  DevBuf feat, cmask, mu_d, x_d, idx, part, res, gate, ngate, in;
  DevBuf odds_d;  // (1 - mu) / mu per sample, at most 1e12 (the Newton passes of the SPA)
  DevBuf limbs, mu2p;  // digit table and zero-padded mu2 of score_sums_mfma_kernel
  SumsMExp fexp{};
  int sums_form = -1;  // uint8 sums kernel: -1 timed choice, else form 0, 1 or 2
  DevBuf zt_coef;      // Chebyshev coefficients of the zero-byte table
  SpaZeroTab zt{};     // intercept-only models (zt.ok): carrier-only byte-table SPA passes
  int spa_split = 1;   // use them for variants with few carriers
  DevBuf spa_vars, spa_part, spa_active;  // batched SPA state
  // sums-kernel shape (variants per lane, sample splits) chosen by timing, per (loader,
  // covariates, variants)
  std::map<std::tuple<int, int, size_t>, std::pair<int, size_t>> tuned_shape;
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

// Device work areas of one batch: the sums partials, the gated-variant list and its count,
// and the batched SPA state.  The allocating entry points take them from the engine's
// buffers (grown on demand); with `fixed` they are given (the scratch array of
// saige_test_markers_u8_device_noalloc) and neither grow nor trigger the shape search.
struct BatchArea {
  bool fixed = false;
  double* part = nullptr;
  size_t part_bytes = 0;
  GateInfo* gate = nullptr;
  size_t gate_cap = 0;  // variants
  uint32_t* ngate = nullptr;
  SpaVar* spa_vars = nullptr;
  size_t spa_vars_cap = 0;  // variants
  double* spa_part = nullptr;
  size_t spa_part_cap = 0;  // doubles
  uint32_t* spa_active = nullptr;  // two counts (256 bytes), then the two variant lists
  size_t spa_list_cap = 0;         // variants per list
};

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
  return std::is_same<L, LoadU8>::value ? 1
       : std::is_same<L, LoadF64>::value ? 2
       : std::is_same<L, LoadPacked>::value ? 3 : 4;
}

// dosages[v * n + i] of a batch of BGEN blocks (the decode of bgen.rs:373-430)
__global__ void bgen_decode_kernel(LoadBgen load, size_t n, size_t nvar,
                                   double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t v = blockIdx.y;
  if (i >= n || v >= nvar) return;
  out[v * n + i] = load(v, i);
}

// SPA of a batch's gated variants with S work-groups per variant and pass (about
// kSpaPassTarget work-groups per pass, slices of at least 2,048 samples: several times the
// work-groups resident at once, so that slices of unequal cost, such as the carrier-only
// slices beside full ones, spread evenly over the CUs); the host reads the number of
// variants still searching after every fourth pass.
template <class G, bool TAB>
static int spa_batched(saige_engine* e, G gt, saige_variant_result* d_out, const BatchArea& ar) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  uint32_t ng = 0;
  SAIGE_HIP(hipMemcpyAsync(&ng, ar.ngate, 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  if (ng == 0) return SAIGE_OK;
  size_t S = std::max<size_t>(1, std::min<size_t>(ceil_div(kSpaPassTarget, ng), ceil_div(n, 2048)));
  S = std::min<size_t>(S, 64);
  const size_t slice = round_up(ceil_div(n, S), 256);
  S = ceil_div(n, slice);
  uint32_t* active;
  SpaVar* vars;
  double* part;
  // partials: ng * S variant-slices for the setup and final passes, at most
  // max(kSpaPassTarget, ng) for a Newton pass; both within kSpaPassTarget + ng
  const size_t nslices = std::max<size_t>(ng * S, kSpaPassTarget + ng);
  if (ar.fixed) {  // ng <= the batch's variants: inside the area
    SAIGE_ARG(ng <= ar.spa_vars_cap && nslices * kSpaPart <= ar.spa_part_cap &&
                  ng <= ar.spa_list_cap,
              "batched SPA state exceeds the scratch area");
    active = ar.spa_active;
    vars = ar.spa_vars;
    part = ar.spa_part;
  } else {
    SAIGE_TRY(e->spa_vars.ensure(ng * sizeof(SpaVar)));
    SAIGE_TRY(e->spa_part.ensure(nslices * kSpaPart * 8));
    SAIGE_TRY(e->spa_active.ensure(256 + 2 * (size_t)ng * 4));
    active = e->spa_active.as<uint32_t>();
    vars = e->spa_vars.as<SpaVar>();
    part = e->spa_part.as<double>();
  }
  GateInfo* gate = ar.gate;
  const uint32_t* ngp = ar.ngate;
  const double* mu = e->mu_d.as<double>();
  const double* odds = e->odds_d.as<double>();
  const MlpView mv = e->mlp_view();
  const dim3 grid(ng, (unsigned)S);
  const unsigned sb = (unsigned)ceil_div(ng, 256);
  const int fast = e->ec.fast, iS = (int)S;
  SpaList sl{active, active + 64, ng, 0};  // counts, then the two halves of the list
  const dim3 grid1((unsigned)std::max<size_t>(kSpaPassTarget, ng));
  auto launch_pass = [&](auto kind) {
    constexpr int KIND = decltype(kind)::value;
    const dim3 gr = KIND == 1 ? grid1 : grid;
    if constexpr (TAB)  // the Newton passes take the odds (1 - mu) / mu in place of mu
      spa_pass_tab_kernel<KIND><<<gr, kSpaPassBlock, 0, s>>>(
          gt.load, KIND == 1 ? odds : mu, n, gate, ngp, fast, slice, vars, part, sl, e->zt);
    else
      spa_pass_kernel<G, KIND><<<gr, kSpaPassBlock, 0, s>>>(gt, mu, n, gate, ngp, fast, slice,
                                                            vars, part, sl);
  };
  zero_u32_kernel<<<1, 1, 0, s>>>(active);
  zero_u32_kernel<<<1, 1, 0, s>>>(active + 1);
  launch_pass(std::integral_constant<int, 0>{});
  const int split_ok = (TAB && e->zt.ok && e->spa_split) ? 1 : 0;
  spa_setup_step<<<(unsigned)ceil_div((size_t)ng * 64, 256), 256, 0, s>>>(gate, ngp, iS, part,
                                                                         vars, d_out, n, split_ok);
  if (mv.enabled) spa_propose_kernel<<<ng, 128, 0, s>>>(gate, ngp, mv, vars);
  spa_start_step<<<sb, 256, 0, s>>>(gate, ngp, vars, sl);
  // A pass with no variant left costs a few microseconds, a read of the count one host
  // round trip: the count is first read after pass 11, then after every fourth pass.
  const unsigned sbw = (unsigned)ceil_div((size_t)ng * 64, 256);
  for (int pass = 0; pass < kSpaMaxPasses; ++pass) {
    sl.pass = pass;
    launch_pass(std::integral_constant<int, 1>{});
    spa_eval_step<<<sbw, 256, 0, s>>>(gate, n, part, e->ec.spa_tol, vars, sl);
    if (pass >= 11 && pass % 4 == 3) {
      uint32_t cnt = 0;
      SAIGE_HIP(hipMemcpyAsync(&cnt, active + ((pass + 1) & 1), 4, hipMemcpyDeviceToHost, s));
      SAIGE_HIP(hipStreamSynchronize(s));
      if (cnt == 0) break;
    }
  }
  launch_pass(std::integral_constant<int, 2>{});
  spa_final_step<<<(unsigned)ceil_div((size_t)ng * 64, 256), 256, 0, s>>>(gate, ngp, iS, part,
                                                                         vars, d_out);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}


// One launch of the sums kernel.  uint8 dosages with 16-byte rows take a byte kernel: form 2
// the int8 matrix cores (score_sums_mfma_kernel), form 0 lanes over samples (up to 3
// covariates), form 1 a lane per variant; the other loaders take the generic kernel.
template <int P>
constexpr bool has_u8t() { return P <= 3; }
constexpr int kSumsMfmaRT = 2;  // 16-variant tiles per wave of score_sums_mfma_kernel
constexpr int kSumsMfmaNW = 8;  // its waves per work-group (sharing the staged table)

template <int P, class L>
static void launch_sums(saige_engine* e, L load, size_t nvar, size_t nvar_pad, int form,
                        size_t splits, size_t sps, double* part) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  if constexpr (std::is_same<L, LoadU8>::value) {
    if (load.vec) {
      if (form == 2) {
        constexpr int RT = kSumsMfmaRT, NW = kSumsMfmaNW;
        score_sums_mfma_kernel<P, RT, NW>
            <<<dim3((unsigned)ceil_div(nvar, 16 * RT * NW), (unsigned)splits), 64 * NW, 0, s>>>(
                load, e->limbs.as<int8_t>(), e->mu2p.as<double>(), e->fexp, round_up(n, 256),
                nvar, sps, part, nvar_pad);
        return;
      }
      const uint32_t* cm = e->cmask.as<uint32_t>();
      if constexpr (has_u8t<P>()) {
        if (form == 0) {
          const size_t vpw = 4 * SumsT<P>::VB;
          score_sums_u8t_kernel<P><<<dim3((unsigned)ceil_div(nvar, vpw), (unsigned)splits), 256, 0,
                                     s>>>(load, e->feat.as<double>(), cm, n, nvar, sps, part,
                                          nvar_pad);
          return;
        }
      }
      score_sums_u8_kernel<P, 1><<<dim3((unsigned)ceil_div(nvar, 64), (unsigned)splits), 64, 0,
                                   s>>>(load, e->feat.as<double>(), cm, n, nvar, sps, part,
                                        nvar_pad);
      return;
    }
  }
  score_sums_kernel<P, L><<<dim3((unsigned)ceil_div(nvar, 64), (unsigned)splits), 64, 0, s>>>(
      load, e->feat.as<double>(), n, nvar, sps, part, nvar_pad);
}

// Sample-range length of a split: whole 256-sample blocks for forms 0 and 2 (at most
// kSumsMaxRange for form 2), 16-sample segments otherwise.
static size_t split_len(size_t n, size_t splits, int form) {
  const size_t g = form == 1 ? 16 : 256;
  const size_t len = round_up(ceil_div(n, std::max<size_t>(splits, 1)), g);
  return form == 2 ? std::min(len, kSumsMaxRange) : len;
}

// Tile-shape autotuner of the sums kernel: the first batch of a shape times the kernel and
// the reduction of its partials for each form and 16 to 256 sample splits (every candidate
// computes the complete, correct sums; the run that follows recomputes them with the
// choice), and the fastest (form, splits) serves the later batches of that shape.  Fewer
// than 64 * 1024 samples use the default shape.
template <int P, class L>
static int sums_shape(saige_engine* e, L load, size_t nvar, size_t nvar_pad, int* form_out,
                      size_t* splits_out, size_t max_splits = 0) {
  constexpr int NS = Layout<P>::NS;
  const size_t n = e->n;
  const bool u8 = std::is_same<L, LoadU8>::value && load.vec;
  const bool mf = u8 && e->limbs.p != nullptr;
  const int form0 = mf ? 2 : (u8 && has_u8t<P>()) ? 0 : 1;
  auto norm = [&](size_t sp, int form) { return ceil_div(n, split_len(n, sp, form)); };
  const size_t vblocks = ceil_div(nvar, 64);
  const size_t dflt_sp = std::max<size_t>(1, std::min<size_t>(ceil_div(2048, vblocks),
                                                               ceil_div(n, 512)));
  const auto key = std::make_tuple(loader_id<L>() * 2 + (int)load.vec, P, nvar);
  if (u8 && e->sums_form >= 0) {  // a fixed form at its default splits
    const int f = e->sums_form;
    size_t sp = norm(f != 1 ? 16 : dflt_sp, f);
    if (max_splits) sp = std::min(sp, max_splits);
    *form_out = f;
    *splits_out = sp;
    return SAIGE_OK;
  }
  auto it = e->tuned_shape.find(key);
  if (it != e->tuned_shape.end() && (max_splits == 0 || it->second.second <= max_splits)) {
    *form_out = it->second.first;
    *splits_out = it->second.second;
    return SAIGE_OK;
  }
  int best_form = form0;
  size_t best = norm(form0 != 1 ? 16 : dflt_sp, form0);
  if (max_splits) {  // no search: the default shape within the split bound
    *form_out = best_form;
    *splits_out = std::min(best, max_splits);
    return SAIGE_OK;
  }
  if (n >= 65536) {
    hipStream_t s = e->stream;
    float best_ms = 1e30f;
    hipEvent_t a, b;
    SAIGE_HIP(hipEventCreate(&a));
    SAIGE_HIP(hipEventCreate(&b));
    std::vector<int> forms;
    if (mf) forms.push_back(2);
    if (u8 && has_u8t<P>()) forms.push_back(0);
    forms.push_back(1);
    for (int form : forms) {
      for (size_t cand : {size_t(16), size_t(32), size_t(64), size_t(128), size_t(256)}) {
        const size_t sp = norm(cand, form);
        const size_t sps = split_len(n, sp, form);
        SAIGE_TRY(e->part.ensure((sp + 1) * NS * nvar_pad * 8));
        double* part = e->part.as<double>();
        SAIGE_HIP(hipEventRecord(a, s));
        launch_sums<P, L>(e, load, nvar, nvar_pad, form, sp, sps, part);
        sums_reduce_kernel<<<dim3((unsigned)ceil_div(nvar_pad, 256), (unsigned)NS), 256, 0, s>>>(
            part, sp, NS, nvar_pad, part + sp * NS * nvar_pad);
        SAIGE_HIP(hipEventRecord(b, s));
        SAIGE_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        SAIGE_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms < best_ms) {
          best_ms = ms;
          best = sp;
          best_form = form;
        }
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  e->tuned_shape[key] = std::make_pair(best_form, best);
  *form_out = best_form;
  *splits_out = best;
  return SAIGE_OK;
}

// Sample splits of the sums kernel at most, in a fixed work area.
constexpr size_t kFixedSplits = 64;

template <int P, class L>
static int run_batch_p(saige_engine* e, L load, size_t nvar, saige_variant_result* d_out,
                       const BatchArea* fixed) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  constexpr int NS = Layout<P>::NS;
  const size_t nvar_pad = round_up(nvar, 128);
  size_t splits = 1;
  int form = 1;
  SAIGE_TRY((sums_shape<P, L>(e, load, nvar, nvar_pad, &form, &splits,
                              fixed ? kFixedSplits : 0)));
  const size_t sps = split_len(n, splits, form);
  SAIGE_ARG(splits * sps >= n, "sample splits cover the samples");
  const size_t part_bytes = (splits + 1) * NS * nvar_pad * 8;
  BatchArea ar;
  if (fixed) {
    ar = *fixed;
    SAIGE_ARG(part_bytes <= ar.part_bytes && nvar <= ar.gate_cap,
              "batch exceeds the scratch area");
  } else {
    SAIGE_TRY(e->part.ensure(part_bytes));
    SAIGE_TRY(e->gate.ensure(nvar * sizeof(GateInfo)));
    SAIGE_TRY(e->ngate.ensure(64));
    ar.part = e->part.as<double>();
    ar.gate = e->gate.as<GateInfo>();
    ar.ngate = e->ngate.as<uint32_t>();
  }
  zero_u32_kernel<<<1, 1, 0, s>>>(ar.ngate);  // a kernel, not a memset node
  launch_sums<P, L>(e, load, nvar, nvar_pad, form, splits, sps, ar.part);
  double* sums = ar.part + splits * NS * nvar_pad;
  sums_reduce_kernel<<<dim3((unsigned)ceil_div(nvar_pad, 256), (unsigned)NS), 256, 0, s>>>(
      ar.part, splits, NS, nvar_pad, sums);
  score_finalize_kernel<P><<<(unsigned)ceil_div(nvar, 128), 128, 0, s>>>(
      sums, 1, nvar, nvar_pad, n, e->ec, L::kMayMiss ? 1 : 0, d_out, ar.gate, ar.ngate);
  SAIGE_HIP(hipGetLastError());
  if (e->ec.use_spa) {
    bool tab = false;
    if constexpr (std::is_same<L, LoadU8>::value) tab = (P == 1 && e->intercept_only);
    if (tab) {
      if constexpr (std::is_same<L, LoadU8>::value)
        SAIGE_TRY((spa_batched<GTildeU8Tab, true>(e, GTildeU8Tab{load}, d_out, ar)));
    } else {
      SAIGE_TRY((spa_batched<GTilde<L, P>, false>(e, GTilde<L, P>{load, e->x_d.as<double>(), n},
                                                   d_out, ar)));
    }
  }
  return SAIGE_OK;
}

template <class L>
static int run_batch(saige_engine* e, L load, size_t nvar, saige_variant_result* d_out,
                     const BatchArea* fixed = nullptr) {
  switch (e->p) {
    case 1: return run_batch_p<1>(e, load, nvar, d_out, fixed);
    case 2: return run_batch_p<2>(e, load, nvar, d_out, fixed);
    case 3: return run_batch_p<3>(e, load, nvar, d_out, fixed);
    case 4: return run_batch_p<4>(e, load, nvar, d_out, fixed);
    case 5: return run_batch_p<5>(e, load, nvar, d_out, fixed);
    case 6: return run_batch_p<6>(e, load, nvar, d_out, fixed);
    case 7: return run_batch_p<7>(e, load, nvar, d_out, fixed);
    case 8: return run_batch_p<8>(e, load, nvar, d_out, fixed);
  }
  set_error("p = %zu covariates not supported (1..8)", e->p);
  return SAIGE_ERR_ARG;
}

constexpr size_t kHostBatch = 4096;

// ---------------------------------------------------------------------------------
// Work area of saige_test_markers_u8_device_noalloc: one file-scope device array (each GPU
// the library runs on holds its own copy until the process ends), carved into one batch's
// work areas, so that the call allocates nothing.  Sized for kScratchVariants variants per
// call at any covariate count (p <= 8) with at most kFixedSplits sample splits of the sums
// kernel.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
// engines on two host threads) race; calls on one engine are serialized by its lock.
// ---------------------------------------------------------------------------------
constexpr size_t kScratchVariants = 4096;
constexpr size_t cround(size_t x, size_t m) { return (x + m - 1) / m * m; }
constexpr size_t kScratchPart = (kFixedSplits + 1) * (2 * kMaxP + 6) * kScratchVariants * 8;
constexpr size_t kScratchGate = cround(kScratchVariants * sizeof(GateInfo), 256);
constexpr size_t kScratchVars = cround(kScratchVariants * sizeof(SpaVar), 256);
// the SPA partials: ng * S <= 1024 + ng variant-slices (S <= ceil(1024 / ng)) in the setup
// and final passes, at most max(kSpaPassTarget, ng) in a Newton pass
constexpr size_t kScratchSpaPart = cround((kSpaPassTarget + kScratchVariants) * kSpaPart * 8, 256);
constexpr size_t kScratchSpaActive = cround(256 + 2 * kScratchVariants * 4, 256);
constexpr size_t kScratchBytes = cround(kScratchPart, 256) + kScratchGate + 256 + kScratchVars +
                                 kScratchSpaPart + kScratchSpaActive;
static_assert(kScratchBytes <= (48u << 20), "scratch stays within 48 MiB");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;

// The batch's work areas in this GPU's scratch array; its address is looked up once per
// device, with that device current, and kept.
static int scratch_area(int device, BatchArea* ar) {
  static std::mutex mu;
  static unsigned char* base_of[kScratchDevices] = {};
  SAIGE_ARG(device >= 0 && device < kScratchDevices, "device index < 64");
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!base_of[device]) {
      void* ptr = nullptr;
      SAIGE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      base_of[device] = static_cast<unsigned char*>(ptr);
    }
    base = base_of[device];
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += cround(bytes, 256);
    return p;
  };
  ar->fixed = true;
  ar->part = reinterpret_cast<double*>(take(kScratchPart));
  ar->part_bytes = kScratchPart;
  ar->gate = reinterpret_cast<GateInfo*>(take(kScratchGate));
  ar->gate_cap = kScratchVariants;
  ar->ngate = reinterpret_cast<uint32_t*>(take(256));
  ar->spa_vars = reinterpret_cast<SpaVar*>(take(kScratchVars));
  ar->spa_vars_cap = kScratchVariants;
  ar->spa_part = reinterpret_cast<double*>(take(kScratchSpaPart));
  ar->spa_part_cap = kScratchSpaPart / 8;
  ar->spa_active = reinterpret_cast<uint32_t*>(take(kScratchSpaActive));
  ar->spa_list_cap = kScratchVariants;
  return off <= kScratchBytes ? SAIGE_OK : SAIGE_ERR_ARG;
}

}  // namespace saige

// The digit table and zero-padded mu2 of score_sums_mfma_kernel: the linear features (A(p),
// res, mu2 x(p)) as NDIG balanced base-128 digits of f 2^-E_f, E_f = ilogb(max |f|) + 2, so
// that |f| 2^-E_f < 1/2 and every digit lies in [-64, 64]; each step takes the nearest
// integer of 128 x and keeps the exact remainder.
static int build_sums_tables(saige_engine* e, const std::vector<double>& feat, size_t nf) {
  const size_t n = e->n, p = e->p, nl = 2 * p + 1;
  constexpr int ND = 10;
  static_assert(ND == SumsM<1>::NDIG, "digits per feature");
  const size_t ncp = round_up(2 + ND * nl, 16), npad = round_up(n, 256);
  auto fcol = [&](size_t f) { return f <= p ? f : f + 1; };  // skips mu2 (column p + 1)
  std::vector<int8_t> tab(npad * ncp, 0);
  std::vector<double> m2(npad, 0.0);
  for (size_t f = 0; f < nl; ++f) {
    double mx = 0.0;
    for (size_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(feat[i * nf + fcol(f)]));
    SAIGE_ARG(std::isfinite(mx), "finite engine inputs");
    e->fexp.e[f] = mx > 0.0 ? std::ilogb(mx) + 2 : 0;
  }
  for (size_t i = 0; i < n; ++i) {
    int8_t* b = tab.data() + (i / 256) * ncp * 256 + (i % 256);
    b[0] = 1;
    b[256] = feat[i * nf + 2 * p + 2] != 0.0 ? 1 : 0;
    for (size_t f = 0; f < nl; ++f) {
      double x = std::ldexp(feat[i * nf + fcol(f)], -e->fexp.e[f]);
      for (int l = 0; l < ND; ++l) {
        const double y = x * 128.0;
        const double d = std::nearbyint(y);
        x = y - d;
        b[(2 + ND * f + l) * 256] = (int8_t)d;
      }
    }
    m2[i] = feat[i * nf + p + 1];
  }
  SAIGE_TRY(e->limbs.ensure(tab.size()));
  SAIGE_TRY(e->mu2p.ensure(npad * 8));
  if (hipMemcpy(e->limbs.p, tab.data(), tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->mu2p.p, m2.data(), npad * 8, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("engine upload failed");
    return SAIGE_ERR_DEVICE;
  }
  return SAIGE_OK;
}

// The zero-byte table (SpaZeroTab) of an intercept-only engine: the exact sums at the
// Chebyshev points of unit intervals covering s in [-ln max rho - 40, -ln min rho + 40], the
// series coefficients by the discrete cosine transform of those values.  Left unbuilt (the
// passes then take every sample) where some mu is within 1e-12 of 0 or 1, or where the
// range exceeds 1,024 intervals.
static int build_spa_zero_table(saige_engine* e, const double* mu) {
  const size_t n = e->n;
  e->zt = SpaZeroTab{};
  if (!e->intercept_only) return SAIGE_OK;
  double rmin = HUGE_VAL, rmax = 0.0;
  SpaZeroTab T{};
  T.n = (double)n;
  for (size_t i = 0; i < n; ++i) {
    if (!(mu[i] > 1e-12 && mu[i] < 1.0 - 1e-12)) return SAIGE_OK;
    const double r = (1.0 - mu[i]) / mu[i];
    rmin = std::min(rmin, r);
    rmax = std::max(rmax, r);
    T.rho_sum += r;
    T.rho_inv_sum += 1.0 / r;
    T.lnmu_sum += std::log(mu[i]);
    T.ln1mmu_sum += std::log1p(-mu[i]);
  }
  T.s_lo = std::floor(-std::log(rmax) - kZtMargin);
  const double hi = -std::log(rmin) + kZtMargin;
  T.npiece = (int)std::ceil(hi - T.s_lo);
  if (T.npiece < 1 || T.npiece > 1024 || T.s_lo < -700.0 || hi > 700.0) return SAIGE_OK;
  T.s_hi = T.s_lo + T.npiece;
  const size_t nn = (size_t)T.npiece * kZtN;
  std::vector<double> nodes(nn), vals(3 * nn);
  const double pi = 3.14159265358979323846;
  for (int p = 0; p < T.npiece; ++p)
    for (int k = 0; k < kZtN; ++k)
      nodes[(size_t)p * kZtN + k] = T.s_lo + p + 0.5 + 0.5 * std::cos(pi * (k + 0.5) / kZtN);
  DevBuf dn, dv;
  SAIGE_TRY(dn.ensure(nn * 8));
  SAIGE_TRY(dv.ensure(3 * nn * 8));
  SAIGE_HIP(hipMemcpy(dn.p, nodes.data(), nn * 8, hipMemcpyHostToDevice));
  spa_zero_nodes_kernel<<<(unsigned)nn, 256, 0, e->stream>>>(
      e->odds_d.as<double>(), e->mu_d.as<double>(), n, dn.as<double>(), dv.as<double>());
  SAIGE_HIP(hipGetLastError());
  SAIGE_HIP(hipMemcpyAsync(vals.data(), dv.p, 3 * nn * 8, hipMemcpyDeviceToHost, e->stream));
  SAIGE_HIP(hipStreamSynchronize(e->stream));
  std::vector<double> coef((size_t)T.npiece * 3 * kZtN);
  for (int p = 0; p < T.npiece; ++p)
    for (int f = 0; f < 3; ++f)
      for (int j = 0; j < kZtN; ++j) {
        double acc = 0.0;
        for (int k = 0; k < kZtN; ++k)
          acc += vals[3 * ((size_t)p * kZtN + k) + f] * std::cos(pi * j * (k + 0.5) / kZtN);
        coef[((size_t)p * 3 + f) * kZtN + j] = 2.0 * acc / kZtN;
      }
  SAIGE_TRY(e->zt_coef.ensure(coef.size() * 8));
  SAIGE_HIP(hipMemcpy(e->zt_coef.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice));
  T.coef = e->zt_coef.as<double>();
  T.ok = 1;
  e->zt = T;
  return SAIGE_OK;
}

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
  const size_t nf = 2 * p + 3;
  std::vector<double> feat(n * nf);
  // case flags as bytes (0/1), zero-padded to whole 16-sample segments
  std::vector<uint8_t> cmask(round_up(n, 16) + 16, 0);
  for (size_t i = 0; i < n; ++i) {
    double* f = &feat[i * nf];
    for (size_t r = 0; r < p; ++r) f[r] = a_flat[i * p + r];  // col-major p x n view
    f[p] = residuals[i];
    f[p + 1] = mu2[i];
    for (size_t j = 0; j < p; ++j) f[p + 2 + j] = mu2[i] * x[j * n + i];
    f[2 * p + 2] = (y && y[i] > 0.5) ? 1.0 : 0.0;
    cmask[i] = (y && y[i] > 0.5) ? 1 : 0;
  }
  int rc = SAIGE_OK;
  auto fail = [&](int code) {
    saige_engine_destroy(e);
    return code;
  };
  if ((rc = e->feat.ensure(feat.size() * 8)) != SAIGE_OK) return fail(rc);
  if ((rc = e->mu_d.ensure(n * 8)) != SAIGE_OK) return fail(rc);
  if ((rc = e->odds_d.ensure(n * 8)) != SAIGE_OK) return fail(rc);
  std::vector<double> odds(n);
  for (size_t i = 0; i < n; ++i) odds[i] = std::min((1.0 - mu[i]) / std::max(mu[i], 1e-300), 1e12);
  if ((rc = e->x_d.ensure(n * p * 8)) != SAIGE_OK) return fail(rc);
  if ((rc = e->cmask.ensure(cmask.size())) != SAIGE_OK) return fail(rc);
  if (hipMemcpy(e->feat.p, feat.data(), feat.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->cmask.p, cmask.data(), cmask.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->mu_d.p, mu, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->odds_d.p, odds.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->x_d.p, x, n * p * 8, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("engine upload failed");
    return fail(SAIGE_ERR_DEVICE);
  }
  if ((rc = build_sums_tables(e, feat, nf)) != SAIGE_OK) return fail(rc);
  if ((rc = build_spa_zero_table(e, mu)) != SAIGE_OK) return fail(rc);
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

int saige_test_markers_u8_device_noalloc(saige_engine* e, const uint8_t* d_q, size_t ldq,
                                         size_t nvar, double scale,
                                         saige_variant_result* d_out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || (d_q && d_out)), "null pointer");
  SAIGE_ARG(ldq >= e->n, "ldq < n");
  if (nvar > kScratchVariants) {
    set_error("invalid argument: saige_test_markers_u8_device_noalloc takes at most %zu "
              "variants per call (the size of its scratch array), got %zu",
              kScratchVariants, nvar);
    return SAIGE_ERR_ARG;
  }
  if (nvar == 0) return SAIGE_OK;
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  BatchArea ar;
  SAIGE_TRY(scratch_area(e->device, &ar));
  const int vec = (ldq % 16 == 0) && ((uintptr_t)d_q % 16 == 0);
  return run_batch(e, LoadU8{d_q, ldq, scale, vec}, nvar, d_out, &ar);
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

// Host-side checks shared by the BGEN entry points.
static int bgen_check(const uint8_t* blocks, size_t ld, size_t nvar, const uint32_t* info,
                      uint32_t n_file, const uint32_t* sample_idx, size_t n) {
  SAIGE_ARG(nvar == 0 || (blocks && info), "null pointer");
  SAIGE_ARG(ld >= 10 + (size_t)n_file, "ld shorter than a layout-2 header");
  if (!sample_idx) SAIGE_ARG(n == n_file, "n must equal the block's N without a subset");
  for (size_t k = 0; sample_idx && k < n; ++k)
    SAIGE_ARG(sample_idx[k] < n_file, "sample_idx out of range");
  for (size_t v = 0; v < nvar; ++v) {
    const uint32_t bits = info[v] & 0xffu;
    if (bits == 0) {  // bgen.rs:396-398
      set_error("BGEN: 0 bits per probability not supported");
      return SAIGE_ERR_ARG;
    }
    SAIGE_ARG(bits <= 32, "bits per probability > 32");
    uint32_t nb;
    std::memcpy(&nb, blocks + v * ld, 4);
    SAIGE_ARG(nb == n_file, "a block's N differs from n_file");
    const size_t need = 10 + (size_t)n_file + ((2 * (size_t)n_file * bits + 7) >> 3);
    SAIGE_ARG(!(info[v] & 0x100u) || need <= ld, "ld shorter than a block's probabilities");
  }
  return SAIGE_OK;
}

int saige_bgen_dosages(int device, const uint8_t* blocks, size_t ld, size_t nvar,
                       const uint32_t* info, uint32_t n_file, const uint32_t* sample_idx,
                       size_t n, double* dosages) {
  clear_error();
  SAIGE_ARG(nvar == 0 || n == 0 || dosages, "null pointer");
  SAIGE_TRY(bgen_check(blocks, ld, nvar, info, n_file, sample_idx, n));
  if (nvar == 0 || n == 0) return SAIGE_OK;
  SAIGE_TRY(use_device(device));
  hipStream_t s;
  SAIGE_HIP(hipStreamCreate(&s));
  int rc = SAIGE_OK;
  {
    DevBuf d_blk, d_info, d_idx, d_out;
    hipError_t err = hipSuccess;
    if ((rc = d_blk.ensure(nvar * ld)) == SAIGE_OK && (rc = d_info.ensure(nvar * 4)) == SAIGE_OK &&
        (rc = d_idx.ensure(std::max<size_t>(n, 1) * 4)) == SAIGE_OK &&
        (rc = d_out.ensure(nvar * n * 8)) == SAIGE_OK) {
      err = hipMemcpyAsync(d_blk.p, blocks, nvar * ld, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) err = hipMemcpyAsync(d_info.p, info, nvar * 4, hipMemcpyHostToDevice, s);
      if (err == hipSuccess && sample_idx)
        err = hipMemcpyAsync(d_idx.p, sample_idx, n * 4, hipMemcpyHostToDevice, s);
      if (err == hipSuccess) {
        LoadBgen load{d_blk.as<uint8_t>(), ld, d_info.as<uint32_t>(),
                      sample_idx ? d_idx.as<uint32_t>() : nullptr, n_file, 0};
        bgen_decode_kernel<<<dim3((unsigned)ceil_div(n, 256), (unsigned)nvar), 256, 0, s>>>(
            load, n, nvar, d_out.as<double>());
        err = hipGetLastError();
      }
      if (err == hipSuccess) err = hipMemcpyAsync(dosages, d_out.p, nvar * n * 8, hipMemcpyDeviceToHost, s);
      if (err == hipSuccess) err = hipStreamSynchronize(s);
      if (err != hipSuccess) {
        set_error("bgen decode: %s", hipGetErrorString(err));
        rc = SAIGE_ERR_DEVICE;
      }
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

int saige_test_markers_bgen(saige_engine* e, const uint8_t* blocks, size_t ld, size_t nvar,
                            const uint32_t* info, uint32_t n_file, const uint32_t* sample_idx,
                            saige_variant_result* out) {
  clear_error();
  SAIGE_ARG(e && (nvar == 0 || out), "null pointer");
  SAIGE_TRY(bgen_check(blocks, ld, nvar, info, n_file, sample_idx, e->n));
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  if (sample_idx) {
    SAIGE_TRY(e->idx.ensure(e->n * 4));
    SAIGE_HIP(hipMemcpyAsync(e->idx.p, sample_idx, e->n * 4, hipMemcpyHostToDevice, s));
  }
  for (size_t v0 = 0; v0 < nvar; v0 += kHostBatch) {
    const size_t nb = std::min(kHostBatch, nvar - v0);
    const size_t blk_bytes = round_up(nb * ld, 256), info_bytes = round_up(nb * 4, 256);
    SAIGE_TRY(e->in.ensure(blk_bytes + info_bytes + nb * sizeof(saige_variant_result)));
    uint8_t* db = e->in.as<uint8_t>();
    uint32_t* di = reinterpret_cast<uint32_t*>(db + blk_bytes);
    saige_variant_result* dr = reinterpret_cast<saige_variant_result*>(db + blk_bytes + info_bytes);
    SAIGE_HIP(hipMemcpyAsync(db, blocks + v0 * ld, nb * ld, hipMemcpyHostToDevice, s));
    SAIGE_HIP(hipMemcpyAsync(di, info + v0, nb * 4, hipMemcpyHostToDevice, s));
    SAIGE_TRY(run_batch(e, LoadBgen{db, ld, di, sample_idx ? e->idx.as<uint32_t>() : nullptr,
                                    n_file, 0},
                        nb, dr));
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

int saige_engine_set_spa_split(saige_engine* e, int enable, int* has_table) {
  clear_error();
  SAIGE_ARG(e, "engine");
  std::lock_guard<std::mutex> lk(e->mu);
  e->spa_split = enable ? 1 : 0;
  if (has_table) *has_table = e->zt.ok;
  return SAIGE_OK;
}

int saige_engine_copy_tuning(saige_engine* dst, saige_engine* src) {
  clear_error();
  SAIGE_ARG(dst && src, "engine");
  SAIGE_ARG(dst->n == src->n && dst->p == src->p, "engines of the same sample and covariate count");
  if (dst == src) return SAIGE_OK;
  std::scoped_lock lk(dst->mu, src->mu);
  dst->tuned_shape = src->tuned_shape;
  dst->sums_form = src->sums_form;
  dst->spa_split = src->spa_split;
  return SAIGE_OK;
}

int saige_engine_set_sums_form(saige_engine* e, int form) {
  clear_error();
  SAIGE_ARG(e, "engine");
  SAIGE_ARG(form >= -1 && form <= 2, "form -1, 0, 1 or 2");
  SAIGE_ARG(form != 0 || e->p <= 3, "form 0 takes at most 3 covariates");
  std::lock_guard<std::mutex> lk(e->mu);
  e->sums_form = form;
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

// ---------------------------------------------------------------------------------
// Region tests: BURDEN, SKAT and SKAT-O (saige-core/src/score_test/region.rs:76-345).
// ---------------------------------------------------------------------------------
// For a region's weighted dosages gw_j = w_j g_j (missing -> 0, region.rs:86-94) and
// g~_j = gw_j - X (A gw_j), the scores S_j = g~_j' res and Phi = tau_e g~' diag(mu2) g~
// (:75-135) come from per-variant sums and one Gram per region:
//   S_j    = w_j (g_j'res - (A g_j)'(X'res))
//   Phi_jk = tau_e w_j w_k (g_j'M2 g_k - (A g_j)'(X'M2 g_k) - (A g_k)'(X'M2 g_j)
//                           + (A g_j)'(X'M2X)(A g_k))
// The per-variant sums are the score-test sums kernel's; the Gram g' diag(mu2) g is
// computed in 64 x 64 tiles (region_gram_kernel), its sample range split over the grid and
// the partials added in split order.
namespace saige {

constexpr int kRegTile = 64;
constexpr int kRegChunk = 32;

struct RegionTile {
  uint32_t rj, rk;  // first variant rows of the tile's j and k blocks
  uint32_t mj, mk;  // their sizes (<= 64)
};

// part[(tile * splits + split) * 4096 + j * 64 + k] = sum over the split's samples of
// g_j mu2 g_k (missing dosages as 0)
__global__ __launch_bounds__(256) void region_gram_kernel(const double* __restrict__ g, size_t ldg,
                                                          const double* __restrict__ feat, int nf,
                                                          int mu2_col, size_t n, size_t sps,
                                                          const RegionTile* __restrict__ tiles,
                                                          double* __restrict__ part) {
  __shared__ double a[kRegTile][kRegChunk + 1], b[kRegTile][kRegChunk + 1];
  const RegionTile t = tiles[blockIdx.x];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const size_t i0 = (size_t)blockIdx.y * sps, i1 = min(i0 + sps, n);
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  for (size_t c0 = i0; c0 < i1; c0 += kRegChunk) {
    __syncthreads();
    for (int e = threadIdx.x; e < kRegTile * kRegChunk; e += 256) {
      const int row = e / kRegChunk, col = e % kRegChunk;
      const size_t i = c0 + col;
      double va = 0.0, vb = 0.0;
      if (i < i1) {
        if ((uint32_t)row < t.mj) {
          const double x = g[(t.rj + row) * ldg + i];
          va = isnan(x) ? 0.0 : x * feat[i * nf + mu2_col];
        }
        if ((uint32_t)row < t.mk) {
          const double x = g[(t.rk + row) * ldg + i];
          vb = isnan(x) ? 0.0 : x;
        }
      }
      a[row][col] = va;
      b[row][col] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int s = 0; s < kRegChunk; ++s) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = a[4 * ty + r][s];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = b[4 * tx + c][s];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(av[r], bv[c], acc[r][c]);
    }
  }
  double* o = part + ((size_t)blockIdx.x * gridDim.y + blockIdx.y) * (kRegTile * kRegTile);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) o[(4 * ty + r) * kRegTile + 4 * tx + c] = acc[r][c];
}

__global__ void region_gram_reduce(const double* __restrict__ part, size_t splits, size_t ntile,
                                   double* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tile = blockIdx.y;
  if (e >= (size_t)kRegTile * kRegTile || tile >= ntile) return;
  double s = 0.0;
  for (size_t q = 0; q < splits; ++q) s += part[(tile * splits + q) * kRegTile * kRegTile + e];
  out[tile * kRegTile * kRegTile + e] = s;
}

// 1 - P(a, x) with P the regularized lower incomplete gamma function: the power series for
// x < a + 1 and the Legendre continued fraction (modified Lentz) otherwise, as in Press et
// al., Numerical Recipes, section 6.2.  This is 1 - ChiSquared::new(2a).cdf(2x) for the
// statrs call of region.rs:174, 230, 323.
static double gamma_p(double a, double x) {
  if (!(x > 0.0)) return 0.0;
  if (std::isinf(x)) return 1.0;
  const double lg = std::lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int k = 0; k < 100000; ++k) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (std::fabs(del) < std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(-x + a * std::log(x) - lg);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int k = 1; k < 100000; ++k) {
    const double an = -k * (k - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - std::exp(-x + a * std::log(x) - lg) * h;
}
static double chi2_sf(double df, double x) { return 1.0 - gamma_p(0.5 * df, 0.5 * x); }

// burden_test, skat_test and skat_o_test (region.rs:143-339) from the scores and Phi
static void region_pvalues(const std::vector<double>& S, const std::vector<double>& phi, size_t m,
                           const double* rho, size_t n_rho, saige_region_result* r) {
  r->pvalue = r->burden_pvalue = r->skat_pvalue = 1.0;
  r->rho = 0.0;
  r->n_markers = (uint32_t)m;
  if (m == 0) return;
  auto P = [&](size_t j, size_t k) { return phi[j * m + k]; };
  double sum_score = 0.0, q_skat = 0.0, sum_phi = 0.0, trace_phi = 0.0, phi_sq = 0.0;
  for (size_t j = 0; j < m; ++j) sum_score += S[j];
  for (size_t j = 0; j < m; ++j) q_skat += S[j] * S[j];
  for (size_t j = 0; j < m; ++j)
    for (size_t k = 0; k < m; ++k) sum_phi += P(j, k);
  for (size_t j = 0; j < m; ++j) trace_phi += P(j, j);
  for (size_t j = 0; j < m; ++j)
    for (size_t k = 0; k < m; ++k) phi_sq += P(j, k) * P(j, k);
  // burden (:157-175)
  if (!(sum_phi < 1e-30)) r->burden_pvalue = chi2_sf(1.0, sum_score * sum_score / sum_phi);
  // SKAT (:191-233)
  if (!(trace_phi < 1e-30)) {
    const double var_q = 2.0 * phi_sq;
    if (!(var_q < 1e-30)) {
      const double kappa = var_q / (2.0 * trace_phi);
      const double df = 2.0 * trace_phi * trace_phi / var_q;
      if (df > 0.0 && kappa > 0.0) r->skat_pvalue = chi2_sf(df, q_skat / kappa);
    }
  }
  // SKAT-O (:246-339)
  const double q_burden = sum_score * sum_score;
  std::vector<double> row_sums(m, 0.0);
  for (size_t j = 0; j < m; ++j)
    for (size_t k = 0; k < m; ++k) row_sums[j] += P(j, k);
  double best = 1.0, best_rho = 0.0;
  for (size_t q = 0; q < n_rho; ++q) {
    const double rh = rho[q];
    const double q_rho = (1.0 - rh) * q_skat + rh * q_burden;
    const double mean_q = (1.0 - rh) * trace_phi + rh * sum_phi;
    if (mean_q < 1e-30) continue;
    double var_q = 0.0;
    for (size_t j = 0; j < m; ++j)
      for (size_t k = 0; k < m; ++k) {
        const double a = (1.0 - rh) * P(j, k) + rh * row_sums[k];
        const double b = (1.0 - rh) * P(k, j) + rh * row_sums[j];
        var_q += a * b;
      }
    var_q *= 2.0;
    if (var_q < 1e-30) continue;
    const double kappa = var_q / (2.0 * mean_q);
    const double df = 2.0 * mean_q * mean_q / var_q;
    if (df > 0.0 && kappa > 0.0) {
      const double pv = chi2_sf(df, q_rho / kappa);
      if (pv < best) {
        best = pv;
        best_rho = rh;
      }
    }
  }
  r->pvalue = best;
  r->rho = best_rho;
}

template <int P>
static int region_batch_p(saige_engine* e, const double* dg, size_t ldg, size_t nvar,
                          const std::vector<size_t>& starts, const double* weights,
                          const double* rho, size_t n_rho, saige_region_result* out) {
  hipStream_t s = e->stream;
  const size_t n = e->n;
  constexpr int NS = Layout<P>::NS, NF = Layout<P>::NF;
  const size_t nreg = starts.size() - 1;
  // per-variant sums: g'A (p), g'res, g'M2 x (p)
  const size_t nvar_pad = round_up(nvar, 128);
  const size_t splits = std::max<size_t>(1, std::min<size_t>(ceil_div(2048, ceil_div(nvar, 64)),
                                                              ceil_div(n, 512)));
  const size_t sps = round_up(ceil_div(n, splits), 16);
  const size_t sp = ceil_div(n, sps);
  SAIGE_TRY(e->part.ensure((sp + 1) * NS * nvar_pad * 8));
  score_sums_kernel<P, LoadF64><<<dim3((unsigned)ceil_div(nvar, 64), (unsigned)sp), 64, 0, s>>>(
      LoadF64{dg, ldg, 1}, e->feat.as<double>(), n, nvar, sps, e->part.as<double>(), nvar_pad);
  double* sums = e->part.as<double>() + sp * NS * nvar_pad;
  sums_reduce_kernel<<<dim3((unsigned)ceil_div(nvar_pad, 256), (unsigned)NS), 256, 0, s>>>(
      e->part.as<double>(), sp, NS, nvar_pad, sums);
  // Gram tiles (k block >= j block) of every region
  std::vector<RegionTile> tiles;
  std::vector<size_t> tile0(nreg + 1, 0);
  for (size_t r = 0; r < nreg; ++r) {
    tile0[r] = tiles.size();
    const size_t a = starts[r], m = starts[r + 1] - starts[r];
    for (size_t j = 0; j < m; j += kRegTile)
      for (size_t k = j; k < m; k += kRegTile)
        tiles.push_back(RegionTile{(uint32_t)(a + j), (uint32_t)(a + k),
                                   (uint32_t)std::min<size_t>(kRegTile, m - j),
                                   (uint32_t)std::min<size_t>(kRegTile, m - k)});
  }
  tile0[nreg] = tiles.size();
  const size_t nt = tiles.size();
  std::vector<double> hs(NS * nvar_pad), hg(nt * kRegTile * kRegTile);
  if (nt > 0) {
    const size_t gsplits = std::max<size_t>(1, std::min<size_t>(ceil_div(2048, nt),
                                                                ceil_div(n, 4096)));
    const size_t gsps = round_up(ceil_div(n, gsplits), kRegChunk);
    const size_t gsp = ceil_div(n, gsps);
    const size_t tile_bytes = round_up(nt * sizeof(RegionTile), 256);
    SAIGE_TRY(e->res.ensure(tile_bytes + (gsp + 1) * nt * kRegTile * kRegTile * 8));
    RegionTile* dt = reinterpret_cast<RegionTile*>(e->res.as<uint8_t>());
    double* gpart = reinterpret_cast<double*>(e->res.as<uint8_t>() + tile_bytes);
    double* gram = gpart + gsp * nt * kRegTile * kRegTile;
    SAIGE_HIP(hipMemcpyAsync(dt, tiles.data(), nt * sizeof(RegionTile), hipMemcpyHostToDevice, s));
    region_gram_kernel<<<dim3((unsigned)nt, (unsigned)gsp), 256, 0, s>>>(
        dg, ldg, e->feat.as<double>(), NF, P + 1, n, gsps, dt, gpart);
    region_gram_reduce<<<dim3((unsigned)ceil_div(kRegTile * kRegTile, 256), (unsigned)nt), 256, 0,
                         s>>>(gpart, gsp, nt, gram);
    SAIGE_HIP(hipGetLastError());
    SAIGE_HIP(hipMemcpyAsync(hg.data(), gram, hg.size() * 8, hipMemcpyDeviceToHost, s));
  }
  SAIGE_HIP(hipMemcpyAsync(hs.data(), sums, hs.size() * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  const EngineConst& ec = e->ec;
  for (size_t r = 0; r < nreg; ++r) {
    const size_t a = starts[r], m = starts[r + 1] - a;
    std::vector<double> S(m), phi(m * m), ag(m * P), xg(m * P), gram(m * m);
    for (size_t t = tile0[r]; t < tile0[r + 1]; ++t) {
      const RegionTile& T = tiles[t];
      for (uint32_t j = 0; j < T.mj; ++j)
        for (uint32_t k = 0; k < T.mk; ++k) {
          const double v = hg[t * kRegTile * kRegTile + j * kRegTile + k];
          const size_t jj = T.rj - a + j, kk = T.rk - a + k;
          gram[jj * m + kk] = v;
          gram[kk * m + jj] = v;  // the tile below the diagonal mirrors this one
        }
    }
    for (size_t j = 0; j < m; ++j) {
      const size_t v = a + j;
      double gres = hs[(2 + P) * nvar_pad + v];
      for (int q = 0; q < P; ++q) {
        ag[j * P + q] = hs[(2 + q) * nvar_pad + v];
        xg[j * P + q] = hs[(4 + P + q) * nvar_pad + v];
        gres -= ag[j * P + q] * ec.xtr[q];
      }
      S[j] = weights[v] * gres;
    }
    for (size_t j = 0; j < m; ++j)
      for (size_t k = 0; k < m; ++k) {
        double v = gram[j * m + k];
        for (int q = 0; q < P; ++q) {
          v -= ag[j * P + q] * xg[k * P + q] + ag[k * P + q] * xg[j * P + q];
          double t = 0.0;
          for (int q2 = 0; q2 < P; ++q2) t += ec.xm2x[q * P + q2] * ag[k * P + q2];
          v += ag[j * P + q] * t;
        }
        phi[j * m + k] = ec.tau_e * weights[a + j] * weights[a + k] * v;
      }
    region_pvalues(S, phi, m, rho, n_rho, &out[r]);
  }
  return SAIGE_OK;
}

}  // namespace saige

extern "C" double saige_chi2_sf(double df, double x) { return saige::chi2_sf(df, x); }

// Groups of whole regions with at most kHostBatch variants (a larger region alone); host
// rows are copied to the engine's workspace per group, device rows are read in place.
static int region_tests_impl(saige_engine* e, const double* g, size_t ldg, bool on_device,
                             size_t nvar, const uint64_t* region_start, size_t n_regions,
                             const double* weights, const double* rho, size_t n_rho,
                             saige_region_result* out) {
  SAIGE_ARG(e && region_start && (n_regions == 0 || out), "null pointer");
  SAIGE_ARG(nvar == 0 || (g && weights), "null pointer");
  SAIGE_ARG(ldg >= e->n, "ldg < n");
  SAIGE_ARG(region_start[0] == 0 && region_start[n_regions] == nvar, "region_start bounds");
  for (size_t r = 0; r < n_regions; ++r)
    SAIGE_ARG(region_start[r + 1] >= region_start[r], "region_start must not decrease");
  static const double kRho[11] = {0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0};
  if (!rho) {  // default_rho_values (region.rs:368-370)
    rho = kRho;
    n_rho = 11;
  }
  std::lock_guard<std::mutex> lk(e->mu);
  SAIGE_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  const size_t n = e->n;
  size_t r0 = 0;
  while (r0 < n_regions) {
    size_t r1 = r0 + 1;
    while (r1 < n_regions && region_start[r1 + 1] - region_start[r0] <= kHostBatch) ++r1;
    const size_t v0 = region_start[r0], nb = region_start[r1] - v0;
    std::vector<size_t> starts(r1 - r0 + 1);
    for (size_t r = r0; r <= r1; ++r) starts[r - r0] = region_start[r] - v0;
    if (nb == 0) {
      for (size_t r = r0; r < r1; ++r) region_pvalues({}, {}, 0, rho, n_rho, &out[r]);
    } else {
      const double* dg = g + v0 * ldg;
      size_t ld = ldg;
      if (!on_device) {
        SAIGE_TRY(e->in.ensure(nb * n * 8));
        SAIGE_HIP(hipMemcpy2DAsync(e->in.p, n * 8, g + v0 * ldg, ldg * 8, n * 8, nb,
                                   hipMemcpyHostToDevice, s));
        dg = e->in.as<double>();
        ld = n;
      }
      int rc = SAIGE_ERR_ARG;
      const double* w = weights + v0;
      switch (e->p) {
        case 1: rc = region_batch_p<1>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 2: rc = region_batch_p<2>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 3: rc = region_batch_p<3>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 4: rc = region_batch_p<4>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 5: rc = region_batch_p<5>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 6: rc = region_batch_p<6>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 7: rc = region_batch_p<7>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
        case 8: rc = region_batch_p<8>(e, dg, ld, nb, starts, w, rho, n_rho, out + r0); break;
      }
      if (rc != SAIGE_OK) return rc;
    }
    r0 = r1;
  }
  return SAIGE_OK;
}

extern "C" int saige_region_tests(saige_engine* e, const double* g, size_t ldg, size_t nvar,
                                  const uint64_t* region_start, size_t n_regions,
                                  const double* weights, const double* rho, size_t n_rho,
                                  saige_region_result* out) {
  clear_error();
  return region_tests_impl(e, g, ldg, false, nvar, region_start, n_regions, weights, rho, n_rho,
                           out);
}

extern "C" int saige_region_tests_device(saige_engine* e, const double* d_g, size_t ldg,
                                         size_t nvar, const uint64_t* region_start,
                                         size_t n_regions, const double* weights,
                                         const double* rho, size_t n_rho,
                                         saige_region_result* out) {
  clear_error();
  return region_tests_impl(e, d_g, ldg, true, nvar, region_start, n_regions, weights, rho, n_rho,
                           out);
}
