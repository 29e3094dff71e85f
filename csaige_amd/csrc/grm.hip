// On-the-fly GRM operator on resident packed genotypes (plain FP64 path).
//
// Reference: OnTheFlyGrm (saige-core/src/glmm/pcg.rs:91-181).  GRM * V = (1/M) Z (Z^T V)
// with z_im = (g_im - 2 af_m) / sqrt(2 af_m (1 - af_m)), missing -> 0 and monomorphic
// markers (sqrt(...) <= 1e-10) -> 0 (pcg.rs:106-115).
//
// Layout in HBM: one SNP-major row of n_pad/4 bytes per marker (n_pad a multiple of 512,
// so a 512-sample stripe of a row is one 128-byte line), samples in model order,
// padding samples and padding markers coded 01 (missing, z = 0).  A 4-entry table per
// marker holds z for the codes 00, 01, 10, 11 in fp64, computed exactly as pcg.rs:107-112.
//
// Stage 1 (T = Z^T V, reduction over samples): one lane per marker, so each lane decodes
// its own row; the 64 lanes of a wave share the sample index, so the row of V^T they read
// is wave-uniform and is served from the scalar cache (one SGPR operand per FMA).
// Stage 2 (Y = Z T, reduction over markers): one lane per sample; the row T[m, :] is again
// wave-uniform.  Both stages split the reduction dimension over the grid for occupancy
// and combine the partial sums in a fixed order, so the product is deterministic.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

namespace saige {

// ---------------------------------------------------------------------------------
// Launch log (see launch_log_pair in common.hpp).
// ---------------------------------------------------------------------------------
namespace {
struct LaunchLog {
  static constexpr size_t kMaxLaunches = 65536;
  struct Rec {
    hipEvent_t a, b;
    uint32_t width, vectors;
    int device;
  };
  struct FreeEvent {
    hipEvent_t e;
    int device;
  };
  std::mutex mu;
  std::atomic<bool> on{false};
  uint32_t only_width = 0;      // record only launches of this width (0: every width)
  std::vector<Rec> recs;        // launches recorded since the last read
  std::vector<FreeEvent> pool;  // events free for reuse, per device
};
LaunchLog& launch_log() {
  static LaunchLog* l = new LaunchLog;  // never destroyed: events outlive static teardown
  return *l;
}
int take_event(LaunchLog& L, int dev, hipEvent_t* e) {
  for (size_t i = 0; i < L.pool.size(); ++i)
    if (L.pool[i].device == dev) {
      *e = L.pool[i].e;
      L.pool.erase(L.pool.begin() + (ptrdiff_t)i);
      return SAIGE_OK;
    }
  SAIGE_HIP(hipEventCreate(e));
  return SAIGE_OK;
}
// Waits for the recorded launches, sums the durations of those of one width (0: all) when
// total_ms is given, and returns every event to the pool.
int drain_launch_log(LaunchLog& L, uint32_t width, double* total_ms, uint64_t* launches) {
  double t = 0.0;
  uint64_t cnt = 0;
  int cur = 0, rc = SAIGE_OK;
  SAIGE_HIP(hipGetDevice(&cur));
  for (const LaunchLog::Rec& r : L.recs) {
    // the pair is reused only after its launch has ended
    const bool done = hipSetDevice(r.device) == hipSuccess && hipEventSynchronize(r.b) == hipSuccess;
    if (total_ms && rc == SAIGE_OK && (width == 0 || r.width == width)) {
      float ms = 0.f;
      if (!done || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
        set_error("launch log: a recorded launch could not be timed");
        rc = SAIGE_ERR_DEVICE;
      }
      t += ms;
      ++cnt;
    }
    L.pool.push_back({r.a, r.device});
    L.pool.push_back({r.b, r.device});
  }
  L.recs.clear();
  (void)hipSetDevice(cur);
  if (total_ms) *total_ms = t;
  if (launches) *launches = cnt;
  return rc;
}
}  // namespace

bool launch_log_on() { return launch_log().on.load(std::memory_order_relaxed); }

std::mutex& launch_log_mutex() { return launch_log().mu; }

int launch_log_pair_locked(LaunchTag tag, hipEvent_t* a, hipEvent_t* b) {
  const uint32_t width = tag.width;
  *a = *b = nullptr;
  LaunchLog& L = launch_log();
  if (!L.on.load(std::memory_order_relaxed) || L.recs.size() >= LaunchLog::kMaxLaunches ||
      (L.only_width != 0 && width != L.only_width))
    return SAIGE_OK;
  int dev = 0;
  SAIGE_HIP(hipGetDevice(&dev));
  hipEvent_t e0, e1;
  SAIGE_TRY(take_event(L, dev, &e0));
  const int rc = take_event(L, dev, &e1);
  if (rc != SAIGE_OK) {
    L.pool.push_back({e0, dev});
    return rc;
  }
  L.recs.push_back({e0, e1, width, tag.vectors, dev});
  *a = e0;
  *b = e1;
  return SAIGE_OK;
}

int upload_repack(hipStream_t s, const uint8_t* bed_body, size_t bpm, size_t n_total,
                  const uint32_t* d_idx, size_t n, const uint32_t* marker_idx, size_t m,
                  uint8_t* d_dst, size_t row_bytes);
int upload_sample_idx(hipStream_t s, const uint32_t* sample_idx, size_t n, size_t n_total,
                      DevBuf& d);
int repack_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                       size_t n, const uint32_t* marker_idx, size_t m, uint8_t* d_dst,
                       size_t row_bytes);
int grm_mixed_prepare(saige_grm* h);
void grm_mixed_release(saige_grm* h);
int grm_mixed_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                      size_t ldo, const double* d_w, const double* tau);

__device__ __forceinline__ double pick(uint32_t code, double z0, double z2, double z3) {
  // code 00 -> z0, 10 -> z2, 11 -> z3, 01 (missing) -> 0
  double hi = (code & 1u) ? z3 : z2;
  double lo = (code & 1u) ? 0.0 : z0;
  return (code & 2u) ? hi : lo;
}

// V (n x k, leading dimension ldv) -> Vt (n_pad x K row-major), zero padded.
template <int K>
__global__ void transpose_v_kernel(const double* __restrict__ v, size_t ldv, size_t n, size_t k,
                                   double* __restrict__ vt, size_t n_pad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
#pragma unroll
  for (int c = 0; c < K; ++c) vt[i * K + c] = (i < n && (size_t)c < k) ? v[i + c * ldv] : 0.0;
}

template <int K>
__global__ __launch_bounds__(256) void grm_s1_f64(const uint8_t* __restrict__ geno,
                                                  size_t row_bytes,
                                                  const double* __restrict__ lut,
                                                  const double* __restrict__ vt,
                                                  size_t bytes_per_split,
                                                  double* __restrict__ tpart, size_t m_pad) {
  const size_t m = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t split = blockIdx.y;
  const size_t b0 = split * bytes_per_split;
  const size_t b1 = min(b0 + bytes_per_split, row_bytes);
  const double z0 = lut[4 * m + 0], z2 = lut[4 * m + 2], z3 = lut[4 * m + 3];
  double acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.0;
  const uint8_t* row = geno + m * row_bytes;
  for (size_t b = b0; b < b1; b += 16) {
    const uint4 w4 = *reinterpret_cast<const uint4*>(row + b);
    const uint32_t words[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t word = words[q];
      const double* vr = vt + (b * 4 + (size_t)q * 16) * K;
#pragma unroll 4
      for (int t = 0; t < 16; ++t) {
        const double z = pick((word >> (2 * t)) & 3u, z0, z2, z3);
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(z, vr[t * K + c], acc[c]);
      }
    }
  }
  double* out = tpart + (split * m_pad + m) * K;
#pragma unroll
  for (int c = 0; c < K; ++c) out[c] = acc[c];
}

// T[m, c] = sum over splits (fixed order).
template <int K>
__global__ void grm_s1_reduce(const double* __restrict__ tpart, size_t splits, size_t m_pad,
                              double* __restrict__ t) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m_pad * K) return;
  double s = 0.0;
  for (size_t q = 0; q < splits; ++q) s += tpart[q * m_pad * K + e];
  t[e] = s;
}

template <int K>
__global__ __launch_bounds__(256) void grm_s2_f64(const uint8_t* __restrict__ geno,
                                                  size_t row_bytes,
                                                  const double* __restrict__ lut,
                                                  const double* __restrict__ t,
                                                  size_t markers_per_split, size_t m,
                                                  double* __restrict__ ypart, size_t n_pad) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t split = blockIdx.y;
  const size_t m0 = split * markers_per_split;
  const size_t m1 = min(m0 + markers_per_split, m);
  const size_t byte = i >> 2;
  const uint32_t shift = 2u * (uint32_t)(i & 3);
  double acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.0;
  size_t mm = m0;
  for (; mm + 4 <= m1; mm += 4) {
    uint32_t codes[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) codes[u] = (geno[(mm + u) * row_bytes + byte] >> shift) & 3u;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double* l = lut + 4 * (mm + u);
      const double z = pick(codes[u], l[0], l[2], l[3]);
      const double* tr = t + (mm + u) * K;
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = fma(z, tr[c], acc[c]);
    }
  }
  for (; mm < m1; ++mm) {
    uint32_t code = (geno[mm * row_bytes + byte] >> shift) & 3u;
    const double* l = lut + 4 * mm;
    const double z = pick(code, l[0], l[2], l[3]);
    const double* tr = t + mm * K;
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = fma(z, tr[c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < K; ++c) ypart[(split * K + c) * n_pad + i] = acc[c];
}

// out[i, c] = (1/M) sum_s ypart[s][c][i]; with w: out = tau0 v / max(w,1e-30) + tau1 * that.
template <int K>
__global__ void grm_s2_reduce(const double* __restrict__ ypart, size_t splits, size_t n_pad,
                              size_t n, size_t k, double inv_m, double* __restrict__ out,
                              size_t ldo, const double* __restrict__ v, size_t ldv,
                              const double* __restrict__ w, double tau0, double tau1) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  double s = 0.0;
  for (size_t q = 0; q < splits; ++q) s += ypart[(q * K + c) * n_pad + i];
  double g = s * inv_m;
  if (w) g = tau0 * v[i + c * ldv] / fmax(w[i], 1e-30) + tau1 * g;
  out[i + c * ldo] = g;
}

__global__ void loco_mask_kernel(double* __restrict__ t, size_t rs, size_t cs, size_t m,
                                 size_t kc, const int* __restrict__ chrom,
                                 const double* __restrict__ scale, LocoCodes lc) {
  const size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (row >= m || c >= kc) return;
  const int cc = lc.code[c];
  if (cc < 0) return;
  double& x = t[row * rs + c * cs];
  x = chrom[row] == cc ? 0.0 : x * scale[cc];
}

bool loco_codes(const saige_grm* h, size_t c0, size_t kc, LocoCodes& lc) {
  if (!h->loco_req || h->loco_n_chrom == 0) return false;
  bool any = false;
  for (size_t c = 0; c < kc; ++c) {
    lc.code[c] = (int16_t)h->loco_req[c0 + c];
    any = any || lc.code[c] >= 0;
  }
  return any;
}

void loco_mask(const saige_grm* h, double* t, size_t rs, size_t cs, size_t m, size_t kc,
               const LocoCodes& lc, hipStream_t s) {
  loco_mask_kernel<<<dim3((unsigned)ceil_div(m, 256), (unsigned)kc), 256, 0, s>>>(
      t, rs, cs, m, kc, h->loco_chrom.as<int>(), h->loco_scale.as<double>(), lc);
}

template <int K>
static int product_k(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                     size_t ldo, const double* d_w, const double* tau, size_t c0) {
  GrmDevice& g = h->g;
  hipStream_t s = h->stream;
  // stage 1 split over samples, stage 2 split over markers, aiming at >= 2048 blocks
  const size_t nblk_m = g.m_pad / 256;
  size_t splits1 = std::min(ceil_div(2048, nblk_m), g.row_bytes / 16);
  splits1 = std::max<size_t>(splits1, 1);
  size_t bps = round_up(ceil_div(g.row_bytes, splits1), 16);
  splits1 = ceil_div(g.row_bytes, bps);
  const size_t nblk_n = g.n_pad / 256;
  size_t splits2 = std::max<size_t>(1, std::min(ceil_div(2048, nblk_n), g.m));
  size_t mps = ceil_div(g.m, splits2);
  splits2 = ceil_div(g.m, mps);

  SAIGE_TRY(h->ws_vt.ensure(g.n_pad * K * 8));
  SAIGE_TRY(h->ws_tpart.ensure(splits1 * g.m_pad * K * 8));
  SAIGE_TRY(h->ws_t.ensure(g.m_pad * K * 8));
  SAIGE_TRY(h->ws_ypart.ensure(splits2 * K * g.n_pad * 8));

  transpose_v_kernel<K><<<(unsigned)ceil_div(g.n_pad, 256), 256, 0, s>>>(
      d_v, ldv, g.n, k, h->ws_vt.as<double>(), g.n_pad);
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)K, (uint32_t)k}, grm_s1_f64<K>, dim3((unsigned)nblk_m, (unsigned)splits1),
                         dim3(256), s, g.geno, g.row_bytes, g.lut, h->ws_vt.as<double>(), bps,
                         h->ws_tpart.as<double>(), g.m_pad));
  grm_s1_reduce<K><<<(unsigned)ceil_div(g.m_pad * K, 256), 256, 0, s>>>(
      h->ws_tpart.as<double>(), splits1, g.m_pad, h->ws_t.as<double>());
  LocoCodes lc;
  if (loco_codes(h, c0, k, lc)) loco_mask(h, h->ws_t.as<double>(), K, 1, g.m, k, lc, s);
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)K, (uint32_t)k}, grm_s2_f64<K>, dim3((unsigned)nblk_n, (unsigned)splits2),
                         dim3(256), s, g.geno, g.row_bytes, g.lut, h->ws_t.as<double>(), mps,
                         g.m, h->ws_ypart.as<double>(), g.n_pad));
  grm_s2_reduce<K><<<dim3((unsigned)ceil_div(g.n, 256), (unsigned)k), 256, 0, s>>>(
      h->ws_ypart.as<double>(), splits2, g.n_pad, g.n, k, 1.0 / (double)g.m, d_out, ldo, d_v,
      ldv, d_w, tau ? tau[0] : 0.0, tau ? tau[1] : 0.0);
  SAIGE_HIP(hipGetLastError());
  h->products += k;
  return SAIGE_OK;
}

// ---------------------------------------------------------------------------------
// SAIGE_PREC_F32: saige-rs's own f32 arithmetic (pcg.rs:128-180 at one rayon thread, where
// the whole marker range is one chunk and the order of every sum is fixed).  v is cast to
// f32; per marker, dot_m = sum_i z_im v_i in sample order starting from 0; per sample,
// acc_i += z_im dot_m in marker order; out = (acc_i * (1/M as f32)) as f64.  Every step is
// one f32 multiply and one f32 add, rounded separately (no contraction), with z_im the f32
// rounding of the fp64 standardized value (pcg.rs:111), so each product is bit-identical
// to the reference's.
// ---------------------------------------------------------------------------------
__global__ void to_f32_kernel(const double* __restrict__ v, size_t ldv, size_t n, size_t k,
                              float* __restrict__ vf, size_t ldf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (i < n && c < k) vf[i + c * ldf] = (float)v[i + c * ldv];
}

// dot[m + c * ldd] = sum over samples in order of z_im * vf_ic; one lane per (marker, column)
__global__ __launch_bounds__(256) void grm_f32_dot_kernel(const uint8_t* __restrict__ geno,
                                                          size_t row_bytes,
                                                          const float* __restrict__ lut32,
                                                          const float* __restrict__ vf, size_t ldf,
                                                          size_t n, size_t m,
                                                          float* __restrict__ dot, size_t ldd) {
#pragma clang fp contract(off)
  const size_t mk = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (mk >= m) return;
  const float z0 = lut32[4 * mk + 0], z1 = lut32[4 * mk + 1], z2 = lut32[4 * mk + 2],
              z3 = lut32[4 * mk + 3];
  const uint8_t* row = geno + mk * row_bytes;
  const float* x = vf + c * ldf;
  float d = 0.0f;
  size_t i = 0;
  for (; i + 16 <= n; i += 16) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(row + i / 4);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const uint32_t code = (w >> (2 * t)) & 3u;
      const float z = code == 0u ? z0 : (code == 1u ? z1 : (code == 2u ? z2 : z3));
      const float prod = z * x[i + t];
      d = d + prod;
    }
  }
  for (; i < n; ++i) {
    const uint32_t code = (row[i / 4] >> (2 * (i % 4))) & 3u;
    const float z = code == 0u ? z0 : (code == 1u ? z1 : (code == 2u ? z2 : z3));
    const float prod = z * x[i];
    d = d + prod;
  }
  dot[mk + c * ldd] = d;
}

// acc_i = sum over markers in order of z_im * dot_mc; out = (acc_i * scale) as f64, or the
// Sigma form tau0 v / max(w, 1e-30) + tau1 * that.  One lane per (sample, column).
__global__ __launch_bounds__(256) void grm_f32_acc_kernel(
    const uint8_t* __restrict__ geno, size_t row_bytes, const float* __restrict__ lut32,
    const float* __restrict__ dot, size_t ldd, size_t n, size_t m, float scale,
    double* __restrict__ out, size_t ldo, const double* __restrict__ v, size_t ldv,
    const double* __restrict__ w, double tau0, double tau1) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (i >= n) return;
  const size_t byte = i >> 2;
  const uint32_t sh = 2u * (uint32_t)(i & 3);
  const float* dc = dot + c * ldd;
  float acc = 0.0f;
  for (size_t mk = 0; mk < m; ++mk) {
    const uint32_t code = (geno[mk * row_bytes + byte] >> sh) & 3u;
    const float z = lut32[4 * mk + code];
    const float prod = z * dc[mk];
    acc = acc + prod;
  }
  const float r = acc * scale;
  double g = (double)r;
  if (w) g = tau0 * v[i + c * ldv] / fmax(w[i], 1e-30) + tau1 * g;
  out[i + c * ldo] = g;
}

static int product_f32(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                       size_t ldo, const double* d_w, const double* tau) {
  GrmDevice& g = h->g;
  hipStream_t s = h->stream;
  SAIGE_ARG(!h->loco_req, "LOCO products are not built for SAIGE_PREC_F32");
  const size_t ldf = round_up(g.n, 64);
  SAIGE_TRY(h->ws_vhi.ensure(ldf * k * 4));
  SAIGE_TRY(h->ws_vlo.ensure(g.m * k * 4));
  float* vf = h->ws_vhi.as<float>();
  float* dot = h->ws_vlo.as<float>();
  to_f32_kernel<<<dim3((unsigned)ceil_div(g.n, 256), (unsigned)k), 256, 0, s>>>(d_v, ldv, g.n, k,
                                                                                 vf, ldf);
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)k, (uint32_t)k}, grm_f32_dot_kernel,
                         dim3((unsigned)ceil_div(g.m, 256), (unsigned)k), dim3(256), s, g.geno,
                         g.row_bytes, h->lut32.as<float>(), vf, ldf, g.n, g.m, dot, g.m));
  const float scale = 1.0f / (float)g.m;
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)k, (uint32_t)k}, grm_f32_acc_kernel,
                         dim3((unsigned)ceil_div(g.n, 256), (unsigned)k), dim3(256), s, g.geno,
                         g.row_bytes, h->lut32.as<float>(), dot, g.m, g.n, g.m, scale, d_out,
                         ldo, d_v, ldv, d_w, tau ? tau[0] : 0.0, tau ? tau[1] : 0.0));
  h->products += k;
  return SAIGE_OK;
}

__global__ void zero_cols_kernel(double* out, size_t ldo, size_t n, size_t k) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i < n && c < k) out[i + c * ldo] = 0.0;
}
__global__ void sigma_diag_only_kernel(double* out, size_t ldo, const double* v, size_t ldv,
                                       const double* w, double tau0, size_t n, size_t k) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i < n && c < k) out[i + c * ldo] = tau0 * v[i + c * ldv] / fmax(w[i], 1e-30) + 0.0;
}

int grm_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out, size_t ldo,
                const double* d_w, const double* tau) {
  GrmDevice& g = h->g;
  if (k == 0) return SAIGE_OK;
  if (g.m == 0) {  // pcg.rs:132-134: no markers -> zero product
    dim3 grid((unsigned)ceil_div(g.n, 256), (unsigned)k);
    if (d_w)
      sigma_diag_only_kernel<<<grid, 256, 0, h->stream>>>(d_out, ldo, d_v, ldv, d_w, tau[0], g.n, k);
    else
      zero_cols_kernel<<<grid, 256, 0, h->stream>>>(d_out, ldo, g.n, k);
    SAIGE_HIP(hipGetLastError());
    return SAIGE_OK;
  }
  if (h->precision == SAIGE_PREC_F32) return product_f32(h, d_v, ldv, k, d_out, ldo, d_w, tau);
  if (h->precision != SAIGE_PREC_FP64) return grm_mixed_product(h, d_v, ldv, k, d_out, ldo, d_w, tau);
  for (size_t c0 = 0; c0 < k; c0 += 64) {
    size_t kc = std::min<size_t>(64, k - c0);
    const double* v = d_v + c0 * ldv;
    double* o = d_out + c0 * ldo;
    int rc;
    if (kc <= 1) rc = product_k<1>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 2) rc = product_k<2>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 4) rc = product_k<4>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 8) rc = product_k<8>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 16) rc = product_k<16>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 24) rc = product_k<24>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 32) rc = product_k<32>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 40) rc = product_k<40>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 48) rc = product_k<48>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else rc = product_k<64>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    if (rc != SAIGE_OK) return rc;
  }
  return SAIGE_OK;
}

// ----------------------------------------------------------------------------------
// Deterministic column dots.
// ----------------------------------------------------------------------------------
__global__ __launch_bounds__(kDotBlock) void dot_pairs_kernel(
    const double* __restrict__ A, size_t lda, const double* __restrict__ B, size_t ldb, size_t na,
    int diag, size_t n, size_t nchunks, double* __restrict__ part) {
  const size_t q = blockIdx.y;
  const size_t a = diag ? q : q % na;
  const size_t b = diag ? q : q / na;
  const double* x = A + a * lda;
  const double* y = B + b * ldb;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk;
  const size_t i1 = min(i0 + (size_t)kDotChunk, n);
  double s = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) s = fma(x[i], y[i], s);
  __shared__ double sh[kDotBlock];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[q * nchunks + blockIdx.x] = sh[0];
}

__global__ void finalize_sums_kernel(const double* __restrict__ part, size_t nchunks, size_t nq,
                                     double* __restrict__ res) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  for (size_t c = 0; c < nchunks; ++c) s += part[q * nchunks + c];
  res[q] = s;
}

// Pinned staging for small results read back on the calling thread (grown on demand and
// kept for the thread's lifetime): the copy is one DMA transfer into host memory.
static void* pinned_staging(size_t bytes) {
  static thread_local void* p = nullptr;
  static thread_local size_t have = 0;
  if (bytes > have) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    have = 0;
    const size_t want = std::max<size_t>(bytes, 64 * 1024);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return nullptr;
    }
    have = want;
  }
  return p;
}

static int dot_pairs(hipStream_t s, const double* A, size_t lda, size_t na, const double* B,
                     size_t ldb, size_t nq, int diag, size_t n, double* h_out, DevBuf& part,
                     DevBuf& res) {
  if (nq == 0) return SAIGE_OK;
  size_t nchunks = std::max<size_t>(1, ceil_div(n, kDotChunk));
  SAIGE_TRY(part.ensure(nq * nchunks * 8));
  SAIGE_TRY(res.ensure(nq * 8));
  dot_pairs_kernel<<<dim3((unsigned)nchunks, (unsigned)nq), kDotBlock, 0, s>>>(
      A, lda, B, ldb, na, diag, n, nchunks, part.as<double>());
  finalize_sums_kernel<<<(unsigned)ceil_div(nq, 256), 256, 0, s>>>(part.as<double>(), nchunks,
                                                                     nq, res.as<double>());
  SAIGE_HIP(hipGetLastError());
  void* pin = pinned_staging(nq * 8);
  if (!pin) {
    set_error("hipHostMalloc failed");
    return SAIGE_ERR_NOMEM;
  }
  SAIGE_HIP(hipMemcpyAsync(pin, res.p, nq * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  memcpy(h_out, pin, nq * 8);
  return SAIGE_OK;
}

int col_dots(hipStream_t s, const double* a, size_t lda, const double* b, size_t ldb, size_t n,
             size_t k, double* h_out, DevBuf& part, DevBuf& res) {
  return dot_pairs(s, a, lda, k, b, ldb, k, 1, n, h_out, part, res);
}

int gram(hipStream_t s, const double* A, size_t lda, size_t na, const double* B, size_t ldb,
         size_t nb, size_t n, double* h_out, DevBuf& part, DevBuf& res) {
  return dot_pairs(s, A, lda, na, B, ldb, na * nb, 0, n, h_out, part, res);
}

int col_dots_dev(hipStream_t s, const double* a, size_t lda, const double* b, size_t ldb,
                 size_t n, size_t k, double* d_out, DevBuf& part) {
  if (k == 0) return SAIGE_OK;
  const size_t nchunks = std::max<size_t>(1, ceil_div(n, kDotChunk));
  SAIGE_TRY(part.ensure(k * nchunks * 8));
  dot_pairs_kernel<<<dim3((unsigned)nchunks, (unsigned)k), kDotBlock, 0, s>>>(
      a, lda, b, ldb, k, 1, n, nchunks, part.as<double>());
  finalize_sums_kernel<<<(unsigned)ceil_div(k, 256), 256, 0, s>>>(part.as<double>(), nchunks, k,
                                                                    d_out);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

int read_doubles(hipStream_t s, const double* d, size_t nq, double* h_out) {
  if (nq == 0) return SAIGE_OK;
  void* pin = pinned_staging(nq * 8);
  if (!pin) {
    set_error("hipHostMalloc failed");
    return SAIGE_ERR_NOMEM;
  }
  SAIGE_HIP(hipMemcpyAsync(pin, d, nq * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  memcpy(h_out, pin, nq * 8);
  return SAIGE_OK;
}

// Per-marker table of standardized values: pcg.rs:106-115 in fp64.
static void build_lut(const double* af, size_t m, size_t m_pad, std::vector<double>& lut,
                      std::vector<double>& sc, std::vector<double>& sh) {
  lut.assign(m_pad * 4, 0.0);
  sc.assign(m_pad, 0.0);
  sh.assign(m_pad, 0.0);
  for (size_t j = 0; j < m; ++j) {
    double p = af[j];
    double denom = std::sqrt(2.0 * p * (1.0 - p));
    if (denom > 1e-10) {
      double mean = 2.0 * p;
      lut[4 * j + 0] = (2.0 - mean) / denom;
      lut[4 * j + 1] = 0.0;  // missing -> mean -> 0
      lut[4 * j + 2] = (1.0 - mean) / denom;
      lut[4 * j + 3] = (0.0 - mean) / denom;
      sc[j] = 1.0 / denom;
      sh[j] = mean / denom;
    }
  }
}

}  // namespace saige

using namespace saige;

namespace saige {

int grm_create_impl(int device, const uint8_t* bed_body, bool on_device, size_t bpm,
                    size_t n_total, const uint32_t* sample_idx, size_t n,
                    const uint32_t* marker_idx, size_t m, const double* af, int precision,
                    saige_grm** out) {
  SAIGE_ARG(out != nullptr, "out");
  *out = nullptr;
  SAIGE_ARG(n > 0, "n > 0");
  SAIGE_ARG(m == 0 || (bed_body && af), "bed_body/af");
  SAIGE_ARG(bpm >= (n_total + 3) / 4, "bytes_per_marker < ceil(n_total/4)");
  SAIGE_ARG(precision == SAIGE_PREC_FP64 || precision == SAIGE_PREC_MIXED ||
                precision == SAIGE_PREC_INT8 || precision == SAIGE_PREC_INT8X2 ||
                precision == SAIGE_PREC_F32, "precision");
  SAIGE_ARG((precision != SAIGE_PREC_INT8 && precision != SAIGE_PREC_INT8X2) ||
                (n <= 8000000 && m <= 8000000),
            "SAIGE_PREC_INT8 and SAIGE_PREC_INT8X2 need n and M <= 8,000,000 (exact int32 "
            "accumulation)");
  if (!sample_idx) SAIGE_ARG(n == n_total, "n must equal n_total without a subset");
  SAIGE_TRY(use_device(device));
  saige_grm* h = new saige_grm();
  h->device = device;
  h->precision = precision;
  auto fail = [&](int rc) {
    saige_grm_destroy(h);
    return rc;
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete h;
    return SAIGE_ERR_DEVICE;
  }
  GrmDevice& g = h->g;
  g.n = n;
  g.m = m;
  g.n_pad = round_up(n, kSamplePad);
  g.m_pad = round_up(std::max<size_t>(m, 1), kMarkerPad);
  g.row_bytes = g.n_pad / 4;
  int rc;
  if ((rc = h->geno.ensure(g.m_pad * g.row_bytes)) != SAIGE_OK) return fail(rc);
  g.geno = h->geno.as<uint8_t>();
  if (hipMemsetAsync(g.geno, 0x55, g.m_pad * g.row_bytes, h->stream) != hipSuccess) {
    set_error("hipMemset failed");
    return fail(SAIGE_ERR_DEVICE);
  }
  DevBuf d_idx;
  if ((rc = upload_sample_idx(h->stream, sample_idx, n, n_total, d_idx)) != SAIGE_OK) return fail(rc);
  if (m > 0) {
    rc = on_device ? repack_from_device(h->stream, bed_body, bpm, d_idx.as<uint32_t>(), n,
                                        marker_idx, m, g.geno, g.row_bytes)
                   : upload_repack(h->stream, bed_body, bpm, n_total, d_idx.as<uint32_t>(), n,
                                   marker_idx, m, g.geno, g.row_bytes);
    if (rc != SAIGE_OK) return fail(rc);
  }
  std::vector<double> lut, sc, sh;
  build_lut(af, m, g.m_pad, lut, sc, sh);
  if (precision == SAIGE_PREC_FP64) {  // the table is read by the FP64 kernels only
    if ((rc = h->lut.ensure(lut.size() * 8)) != SAIGE_OK) return fail(rc);
    g.lut = h->lut.as<double>();
    if (hipMemcpyAsync(g.lut, lut.data(), lut.size() * 8, hipMemcpyHostToDevice, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      set_error("lut upload failed");
      return fail(SAIGE_ERR_DEVICE);
    }
  } else if (precision == SAIGE_PREC_F32) {  // f32 rounding of each standardized value
    std::vector<float> l32(lut.size());
    for (size_t e = 0; e < lut.size(); ++e) l32[e] = (float)lut[e];
    if ((rc = h->lut32.ensure(l32.size() * 4)) != SAIGE_OK) return fail(rc);
    if (hipMemcpyAsync(h->lut32.p, l32.data(), l32.size() * 4, hipMemcpyHostToDevice,
                       h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      set_error("lut upload failed");
      return fail(SAIGE_ERR_DEVICE);
    }
  } else {
    h->host_scale.swap(sc);
    h->host_shift.swap(sh);
    if ((rc = grm_mixed_prepare(h)) != SAIGE_OK) return fail(rc);
  }
  *out = h;
  return SAIGE_OK;
}

}  // namespace saige

extern "C" {

int saige_launch_log_enable(int enable, uint32_t width) {
  using namespace saige;
  clear_error();
  LaunchLog& L = launch_log();
  std::lock_guard<std::mutex> lk(L.mu);
  SAIGE_TRY(drain_launch_log(L, 0, nullptr, nullptr));
  L.only_width = width;
  L.on.store(enable != 0, std::memory_order_relaxed);
  return SAIGE_OK;
}

int saige_launch_log_summary(uint32_t* widths, double* total_ms, uint64_t* launches,
                             uint64_t* vectors, size_t cap, size_t* n_widths) {
  using namespace saige;
  clear_error();
  SAIGE_ARG(n_widths && (cap == 0 || (widths && total_ms && launches && vectors)),
            "null pointer");
  LaunchLog& L = launch_log();
  std::lock_guard<std::mutex> lk(L.mu);
  struct Agg {
    double ms = 0.0;
    uint64_t count = 0, vec = 0;
  };
  std::vector<std::pair<uint32_t, Agg>> agg;  // in increasing width
  int cur = 0, rc = SAIGE_OK;
  SAIGE_HIP(hipGetDevice(&cur));
  for (const LaunchLog::Rec& r : L.recs) {
    float ms = 0.f;
    const bool ok = hipSetDevice(r.device) == hipSuccess && hipEventSynchronize(r.b) == hipSuccess &&
                    hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess;
    if (!ok && rc == SAIGE_OK) {
      set_error("launch log: a recorded launch could not be timed");
      rc = SAIGE_ERR_DEVICE;
    }
    auto it = std::lower_bound(agg.begin(), agg.end(), r.width,
                               [](const std::pair<uint32_t, Agg>& x, uint32_t w) { return x.first < w; });
    if (it == agg.end() || it->first != r.width) it = agg.insert(it, {r.width, Agg{}});
    it->second.ms += ms;
    it->second.count += 1;
    it->second.vec += r.vectors;
    L.pool.push_back({r.a, r.device});
    L.pool.push_back({r.b, r.device});
  }
  L.recs.clear();
  (void)hipSetDevice(cur);
  *n_widths = agg.size();
  for (size_t i = 0; i < agg.size() && i < cap; ++i) {
    widths[i] = agg[i].first;
    total_ms[i] = agg[i].second.ms;
    launches[i] = agg[i].second.count;
    vectors[i] = agg[i].second.vec;
  }
  return rc;
}

int saige_launch_log_read(uint32_t width, double* total_ms, uint64_t* launches) {
  using namespace saige;
  clear_error();
  SAIGE_ARG(total_ms && launches, "total_ms and launches");
  LaunchLog& L = launch_log();
  std::lock_guard<std::mutex> lk(L.mu);
  return drain_launch_log(L, width, total_ms, launches);
}

int saige_grm_create(int device, const uint8_t* bed_body, size_t bpm, size_t n_total,
                     const uint32_t* sample_idx, size_t n, const uint32_t* marker_idx, size_t m,
                     const double* af, int precision, saige_grm** out) {
  clear_error();
  return grm_create_impl(device, bed_body, false, bpm, n_total, sample_idx, n, marker_idx, m, af,
                         precision, out);
}

int saige_grm_create_device(int device, const uint8_t* d_bed_body, size_t bpm, size_t n_total,
                            const uint32_t* sample_idx, size_t n, const uint32_t* marker_idx,
                            size_t m, const double* af, int precision, saige_grm** out) {
  clear_error();
  return grm_create_impl(device, d_bed_body, true, bpm, n_total, sample_idx, n, marker_idx, m, af,
                         precision, out);
}

void saige_grm_destroy(saige_grm* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  grm_mixed_release(h);
  delete h;
}

size_t saige_grm_n_samples(const saige_grm* h) { return h ? h->g.n : 0; }
size_t saige_grm_n_markers(const saige_grm* h) { return h ? h->g.m : 0; }
void* saige_grm_stream(saige_grm* h) { return h ? (void*)h->stream : nullptr; }

int saige_grm_set_kernel_timing(saige_grm* h, int enable) {
  clear_error();
  SAIGE_ARG(h, "handle");
  SAIGE_HIP(hipSetDevice(h->device));
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_TRY(h->timer.read(nullptr, nullptr));
  h->timer.on = enable != 0;
  return SAIGE_OK;
}

int saige_grm_kernel_timing(saige_grm* h, double* total_ms, uint64_t* launches) {
  clear_error();
  SAIGE_ARG(h, "handle");
  SAIGE_HIP(hipSetDevice(h->device));
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  return h->timer.read(total_ms, launches);
}

int saige_grm_synchronize(saige_grm* h) {
  clear_error();
  SAIGE_ARG(h, "handle");
  SAIGE_HIP(hipSetDevice(h->device));
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  return SAIGE_OK;
}

int saige_grm_mat_mat_device(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                             size_t ldo) {
  clear_error();
  SAIGE_ARG(h && (k == 0 || (d_v && d_out)), "null pointer");
  SAIGE_ARG(ldv >= h->g.n && ldo >= h->g.n, "leading dimension < n");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  return grm_product(h, d_v, ldv, k, d_out, ldo);
}

int saige_grm_mat_mat(saige_grm* h, const double* v, size_t ldv, size_t k, double* out,
                      size_t ldo) {
  clear_error();
  SAIGE_ARG(h && (k == 0 || (v && out)), "null pointer");
  SAIGE_ARG(ldv >= h->g.n && ldo >= h->g.n, "leading dimension < n");
  if (k == 0) return SAIGE_OK;
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  const size_t n = h->g.n;
  SAIGE_TRY(h->ws_in.ensure(n * k * 8));
  SAIGE_TRY(h->ws_out.ensure(n * k * 8));
  SAIGE_HIP(hipMemcpy2DAsync(h->ws_in.p, n * 8, v, ldv * 8, n * 8, k, hipMemcpyHostToDevice,
                             h->stream));
  SAIGE_TRY(grm_product(h, h->ws_in.as<double>(), n, k, h->ws_out.as<double>(), n));
  SAIGE_HIP(hipMemcpy2DAsync(out, ldo * 8, h->ws_out.p, n * 8, n * 8, k, hipMemcpyDeviceToHost,
                             h->stream));
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  return SAIGE_OK;
}

int saige_grm_mat_vec(saige_grm* h, const double* v, double* out) {
  return saige_grm_mat_mat(h, v, h ? h->g.n : 0, 1, out, h ? h->g.n : 0);
}

int saige_grm_set_marker_chromosomes(saige_grm* h, const int32_t* chrom) {
  clear_error();
  SAIGE_ARG(h, "null pointer");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  const size_t m = h->g.m;
  if (!chrom) {
    h->loco_n_chrom = 0;
    return SAIGE_OK;
  }
  int nc = 0;
  for (size_t j = 0; j < m; ++j) {
    SAIGE_ARG(chrom[j] >= 0 && chrom[j] < 4096, "chromosome codes in [0, 4096)");
    nc = std::max(nc, chrom[j] + 1);
  }
  std::vector<size_t> count(nc, 0);
  for (size_t j = 0; j < m; ++j) ++count[chrom[j]];
  // dense codes: run_loco fits only the chromosomes present (loco.rs:62-95), so a code
  // without markers would be a refit of a chromosome that does not exist
  for (int c = 0; c < nc; ++c)
    if (count[c] == 0) {
      set_error("invalid argument: chromosome code %d has no marker (codes must be 0..C-1, "
                "each used)", c);
      return SAIGE_ERR_ARG;
    }
  std::vector<double> scale(std::max(nc, 1), 1.0);
  for (int c = 0; c < nc; ++c)  // M / M_c; a GRM without markers is the zero product
    scale[c] = count[c] == m ? 0.0 : (double)m / (double)(m - count[c]);
  SAIGE_TRY(h->loco_chrom.ensure(std::max<size_t>(m, 1) * 4));
  SAIGE_TRY(h->loco_scale.ensure(scale.size() * 8));
  SAIGE_HIP(hipMemcpyAsync(h->loco_chrom.p, chrom, m * 4, hipMemcpyHostToDevice, h->stream));
  SAIGE_HIP(hipMemcpyAsync(h->loco_scale.p, scale.data(), scale.size() * 8,
                           hipMemcpyHostToDevice, h->stream));
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  h->loco_n_chrom = nc;
  return SAIGE_OK;
}

int saige_grm_mat_mat_loco(saige_grm* h, const double* v, size_t ldv, size_t k,
                           int32_t exclude_chrom, double* out, size_t ldo) {
  clear_error();
  SAIGE_ARG(h, "null pointer");
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_ARG(h->loco_n_chrom > 0, "no marker chromosomes set (saige_grm_set_marker_chromosomes)");
  SAIGE_ARG(exclude_chrom >= 0 && exclude_chrom < h->loco_n_chrom, "chromosome code");
  std::vector<int> codes(k, exclude_chrom);
  h->loco_req = codes.data();
  const int rc = saige_grm_mat_mat(h, v, ldv, k, out, ldo);
  h->loco_req = nullptr;
  return rc;
}

}  // extern "C"
