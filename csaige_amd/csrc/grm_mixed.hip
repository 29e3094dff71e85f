// Mixed-precision GRM product: exact genotypes on bf16 MFMA (gfx950), standardization
// factored out, fp32 accumulation, fp64 epilogues.
//
// Same operator as OnTheFlyGrm::mat_vec (saige-core/src/glmm/pcg.rs:128-180):
//   GRM V = (1/M) Z Z^T V,  z_im = (g_im - 2 p_m) s_m  with s_m = 1/sqrt(2 p_m (1 - p_m)).
// Factored:  Z^T V = diag(s) (G^T V - 2p * 1^T V)  and  Z W = G (s*W) - 1 (2p)^T (s*W),
// so the only n x M work is the two integer-matrix products G^T V and G W', run as bf16
// MFMA with G in {0, 1, 2} exact and the dense operand split into bf16 hi + lo.
//
// Conditioning: the dense operand is centred per column before the split (V' = V - mean,
// W' = W - mean over markers) so that the fp32 MFMA accumulators never carry the large
// common component that the factored form subtracts again; the means re-enter exactly
// through fp64 rank-one terms:
//   T_mc   = s_m (U'_mc - 2 p_m cs'_c) + zs_m mean_c,     U' = G^T V'_split,
//            cs'_c = sum_i V'_split_ic,  zs_m = sum_i z_im (0 up to rounding)
//   W_mc   = s_m T_mc;  W' = W - nu_c (nu_c the mean over real markers)
//   M Y_ic = (G W'_split)_ic - sum_m 2 p_m W'_split_mc + nu_c (r_i - P2)
//            with r_i = sum_m g_im and P2 = sum_m 2 p_m.
// Missing genotypes (code 01) decode to 0 in the MFMA operand; sparse per-marker and
// per-sample corrections restore saige-rs's mean imputation (see "Missing calls" below).
//
// Layouts: G is resident twice, SNP-major (stage 1 reads a marker's samples) and
// sample-major (stage 2 reads a sample's markers), 2 bits per genotype in both.  One
// MFMA kernel serves both stages: rows (markers, then samples) x 16*KT columns, reducing
// over the other dimension in 32-long steps (v_mfma_f32_16x16x32_bf16).
#include "common.hpp"

#include <algorithm>
#include <cmath>

namespace saige {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

int marker_stats_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                             size_t n, size_t m, double* ac, uint32_t* n_valid, double* af,
                             double* mac);
int marker_ac_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, size_t n, size_t m,
                          double* d_ac);

// zs_m = s_m (ac_m - 2 p_m n) for real markers, 0 for the padding (rounded as written: no
// contraction into a fused multiply-add)
__global__ void marker_zs_kernel(const double* __restrict__ sc, const double* __restrict__ p2,
                                 const double* __restrict__ ac, double n, size_t m, size_t m_pad,
                                 double* __restrict__ zs) {
#pragma clang fp contract(off)
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m_pad) return;
  zs[j] = j < m ? sc[j] * (ac[j] - p2[j] * n) : 0.0;
}

constexpr int kMixRT = 8;                     // 16-row MFMA tiles per wave
constexpr int kMixRows = 4 * 16 * kMixRT;      // rows per work-group (4 waves)

// ---------------------------------------------------------------------------------
// Prepare: sample-major copy, per-sample row sums, missing check.
// ---------------------------------------------------------------------------------
// A 4 x 4 matrix of 2-bit fields (byte r = row, field c = column) transposed in place: the
// fields of the 2 x 2 sub-blocks swap first, then the two off-diagonal sub-blocks.
__device__ __forceinline__ uint32_t transpose_2bit_4x4(uint32_t w) {
  uint32_t t = (w ^ (w >> 6)) & 0x00CC00CCu;
  w ^= t ^ (t << 6);
  t = (w ^ (w >> 12)) & 0x0000F0F0u;
  w ^= t ^ (t << 12);
  return w;
}

// o[e] = byte e of d0 | byte e of d1 << 8 | byte e of d2 << 16 | byte e of d3 << 24
__device__ __forceinline__ void transpose_bytes_4x4(uint32_t d0, uint32_t d1, uint32_t d2,
                                                    uint32_t d3, uint32_t* o) {
  const uint32_t a = __builtin_amdgcn_perm(d1, d0, 0x05010400u);
  const uint32_t b = __builtin_amdgcn_perm(d1, d0, 0x07030602u);
  const uint32_t c = __builtin_amdgcn_perm(d3, d2, 0x05010400u);
  const uint32_t d = __builtin_amdgcn_perm(d3, d2, 0x07030602u);
  o[0] = __builtin_amdgcn_perm(c, a, 0x05040100u);
  o[1] = __builtin_amdgcn_perm(c, a, 0x07060302u);
  o[2] = __builtin_amdgcn_perm(d, b, 0x05040100u);
  o[3] = __builtin_amdgcn_perm(d, b, 0x07060302u);
}

constexpr int kTrStride = 68;  // bytes per staged marker row (64 used)

// geno: m_pad rows of rb bytes (n_pad/4); geno_t: n_pad rows of rbt bytes (m_pad/4).
// A work-group covers 256 samples and loops over `tiles_per_group` tiles of 256 markers:
// the tile's 64-byte row segments go to LDS, then each thread turns 16 markers x 4 samples
// into 4 output words (four 4 x 4 bit transposes and one byte transpose) and stores them,
// a wave writing 64 contiguous bytes of 4 sample rows per store.  The dosage sum of each
// sample over real markers (code 00 -> 2, 10 -> 1: 2 #{lo = 0} - #{lo = 0, hi = 1}) and the
// number of missing calls (code 01) come from popcounts of the output words; each sample's
// sum reaches global memory once per work-group (integer atomics, order-free).
__global__ __launch_bounds__(256) void transpose_codes_kernel(
    const uint8_t* __restrict__ geno, size_t rb, uint8_t* __restrict__ geno_t, size_t rbt,
    size_t n, size_t m, size_t m_pad, size_t tiles_per_group, unsigned* __restrict__ rsum,
    unsigned* __restrict__ n_miss) {
  __shared__ __attribute__((aligned(16))) uint8_t tile[256 * kTrStride];
  __shared__ unsigned red[256];
  const int t = threadIdx.x;
  const size_t s0 = (size_t)blockIdx.y * 256;
  const int mb = t & 15, sbl = t >> 4;  // markers 16 mb..16 mb + 15, sample bytes sbl + 16 j
  unsigned sum[16], miss = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) sum[i] = 0;
  const size_t tile0 = (size_t)blockIdx.x * tiles_per_group;
  const size_t tile1 = min(tile0 + tiles_per_group, m_pad / 256);
  for (size_t tl = tile0; tl < tile1; ++tl) {
    const size_t m0 = tl * 256;
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // row (t >> 2) + 64 u, 16 bytes at 16 (t & 3)
      const int row = (t >> 2) + 64 * u, seg = t & 3;
      const uint4 v =
          *reinterpret_cast<const uint4*>(geno + (m0 + row) * rb + (s0 >> 2) + 16 * seg);
      uint32_t* d = reinterpret_cast<uint32_t*>(tile + row * kTrStride + 16 * seg);
      d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
    __syncthreads();
    const long long real = (long long)m - (long long)(m0 + 16 * mb);
    const uint32_t rmask =
        real >= 16 ? 0xFFFFFFFFu : (real <= 0 ? 0u : ((1u << (2 * real)) - 1u));
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sb = sbl + 16 * j;
      uint32_t q[4];  // markers 4 i + r (byte r) of 4 samples (fields) -> byte = sample
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        uint32_t w = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          w |= (uint32_t)tile[(16 * mb + 4 * i + r) * kTrStride + sb] << (8 * r);
        q[i] = transpose_2bit_4x4(w);
      }
      uint32_t o[4];  // sample 4 sb + u: its 16 markers in order
      transpose_bytes_4x4(q[0], q[1], q[2], q[3], o);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const size_t i = s0 + 4 * sb + u;
        *reinterpret_cast<uint32_t*>(geno_t + i * rbt + (m0 >> 2) + 4 * mb) = o[u];
        const uint32_t lo = o[u] & 0x55555555u & rmask, hi = (o[u] >> 1) & 0x55555555u & rmask;
        const uint32_t lo0 = ~o[u] & 0x55555555u & rmask;
        sum[4 * j + u] += 2u * __popc(lo0) - __popc(lo0 & hi);
        if (i < n) miss += __popc(lo & ~hi);
      }
    }
    __syncthreads();  // the tile is refilled next
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      red[t] = sum[4 * j + u];
      __syncthreads();
      if (mb == 0) {
        unsigned acc = 0;
        for (int b = 0; b < 16; ++b) acc += red[t + b];
        const size_t i = s0 + 4 * (sbl + 16 * j) + u;
        if (i < n && acc) atomicAdd(&rsum[i], acc);
      }
      __syncthreads();
    }
  red[t] = miss;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0 && red[0]) atomicAdd(n_miss, red[0]);
}

// The int8 GEMM streams the genotypes in blocks of 512 rows x kI8Chunk reduction indices,
// each block 16 KB contiguous, so that a work-group's whole reduction range is one run of
// HBM and every block lands in LDS unchanged by 16-byte DMA loads.  Block (row / 512,
// red / kI8Chunk) starts at ((row / 512) * (red_len / kI8Chunk) + red / kI8Chunk) * 16384;
// inside it the 32-bit words (16 reduction indices of one row) are ordered
// [step s = (red % 128) / 64][row group rg = (row % 512) / 64][lane 16 q + r][tile g]
// with r = row % 16, g = (row % 64) / 16, q = (red % 64) / 16: one 16-byte LDS read gives a
// lane its words of four 16-row tiles for one 64-long MFMA step.  The words hold genotype
// values, not PLINK codes: each 2-bit field is the allele count 0, 1 or 2 (codes 00 -> 2,
// 10 -> 1, 11 -> 0, and the missing code 01 -> 0), so that (w >> 2u) & 0x03030303 is an
// MFMA operand word.
constexpr size_t kI8Chunk = 128;
constexpr size_t kI8Block = 512 * kI8Chunk / 4;  // bytes per (row block, chunk)

__device__ __forceinline__ uint32_t codes_to_values(uint32_t w) {
  const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
  const uint32_t nlo = ~lo & 0x55555555u;  // set for codes 00 and 10
  return ((nlo & ~hi & 0x55555555u) << 1) | (nlo & hi);
}

// One work-group per (row block rb, four chunks kq): the 512 rows' 128-byte runs of the four
// chunks are read as whole lines (eight threads per row), converted to values and placed in
// LDS in the blocked order, then written out as the four contiguous 16 KB blocks.
__global__ __launch_bounds__(256) void block_codes_kernel(const uint8_t* __restrict__ src,
                                                          size_t row_bytes,
                                                          uint8_t* __restrict__ dst) {
  constexpr int kUnits = 4 * (int)kI8Block / 16;  // 16-byte units in four blocks
  __shared__ __attribute__((aligned(16))) uint32_t sh[4 * kI8Block / 4];
  const size_t kq = blockIdx.x, rb = blockIdx.y;
  const size_t nkc = row_bytes * 4 / kI8Chunk;
  for (int i = 0; i < kUnits / 256; ++i) {
    const int u = (int)threadIdx.x + 256 * i;  // row u / 8, 16-byte unit u % 8 of its 128 bytes
    const int row = u >> 3, j = u & 7;
    const uint4 x = *reinterpret_cast<const uint4*>(src + (rb * 512 + row) * row_bytes +
                                                    kq * 128 + 16 * j);
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
    const int rg = row >> 6, g = (row >> 4) & 3, r = row & 15;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int kd = 4 * j + d;  // 16-index word of the 512: chunk kd / 8, step, quarter
      const int kcl = kd >> 3, s = (kd >> 2) & 1, q = kd & 3;
      sh[(kcl * (int)kI8Block + ((s * 8 + rg) * 64 + 16 * q + r) * 16 + g * 4) / 4] =
          codes_to_values(w[d]);
    }
  }
  __syncthreads();
  uint4* out = reinterpret_cast<uint4*>(dst + (rb * nkc + 4 * kq) * kI8Block);
  for (int i = 0; i < kUnits / 256; ++i) {
    const int u = (int)threadIdx.x + 256 * i;
    out[u] = reinterpret_cast<const uint4*>(sh)[u];
  }
}

// buf (rows x row_bytes of PLINK codes, row-major) replaced by its blocked value form
static int block_codes(hipStream_t s, DevBuf& buf, size_t rows, size_t row_bytes) {
  SAIGE_ARG(rows % 512 == 0 && row_bytes % 128 == 0, "blocked genotype layout needs 512-row "
            "blocks and 512-index reduction groups");
  DevBuf tmp;
  SAIGE_TRY(tmp.ensure(rows * row_bytes));
  block_codes_kernel<<<dim3((unsigned)(row_bytes / 128), (unsigned)(rows / 512)), 256, 0, s>>>(
      buf.as<uint8_t>(), row_bytes, tmp.as<uint8_t>());
  SAIGE_HIP(hipGetLastError());
  SAIGE_HIP(hipStreamSynchronize(s));  // the row-major buffer goes back to the cache
  std::swap(buf.p, tmp.p);
  std::swap(buf.bytes, tmp.bytes);
  std::swap(buf.device, tmp.device);
  return SAIGE_OK;
}

__global__ void u32_to_f64_kernel(const unsigned* __restrict__ a, size_t n, double* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)a[i];
}

// ---------------------------------------------------------------------------------
// MFMA kernel: part[split][row][16*KT] = sum over the split's reduction range of
//   code(row, red) * (hi + lo)[col][red]
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bf16x8 decode8(uint32_t bits16, const uint32_t* lut) {
  // 8 two-bit codes -> 8 bf16 dosages (00 -> 2, 10 -> 1, 11 -> 0, 01 missing -> 0)
  union { uint32_t u[4]; bf16x8 v; } o;
#pragma unroll
  for (int p = 0; p < 4; ++p) o.u[p] = lut[(bits16 >> (4 * p)) & 15u];
  return o.v;
}

// Per chunk of KC reduction indices the work-group copies its 512
// code rows (KC/4 bytes each) and the KP x KC hi/lo operand tile into LDS with 16-byte
// coalesced loads (the next chunk is prefetched into registers while the MFMAs of the
// current one run).  Each wave holds its B fragments of one 32-step in registers and
// reuses them for its 8 row tiles.  Strides are padded so that the 2-byte code reads
// (rows at 40 bytes) and the 16-byte operand reads (columns at 272 bytes) hit distinct
// LDS banks.
constexpr int kKC = 128;
constexpr int kArow = 40;              // bytes per staged code row (32 used)
constexpr int kBcol = kKC + 8;         // bf16 per staged operand column (kKC used)

template <int KT>
__global__ __launch_bounds__(256) void mix_gemm_lds_kernel(
    const uint8_t* __restrict__ codes, size_t code_stride, const __bf16* __restrict__ bhi,
    const __bf16* __restrict__ blo, size_t b_stride, size_t red_per_split, size_t red_len,
    float* __restrict__ part, size_t rows_pad) {
  constexpr int KP = 16 * KT;
  constexpr int A_LOADS = kMixRows * (kKC / 4) / 16 / 256;      // 16-byte loads per thread
  constexpr int B_SEGS = kKC * 2 / 16;                           // 16-byte segments per column
  constexpr int B_LOADS = (2 * KP * B_SEGS + 255) / 256;         // per thread
  __shared__ uint32_t lut[16];
  __shared__ __attribute__((aligned(16))) uint8_t sA[kMixRows * kArow];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2 * KP * kBcol];
  if (threadIdx.x < 16) {
    const uint32_t c0 = threadIdx.x & 3u, c1 = threadIdx.x >> 2;
    auto val = [](uint32_t c) -> uint32_t { return c == 0u ? 0x4000u : (c == 2u ? 0x3F80u : 0u); };
    lut[threadIdx.x] = val(c0) | (val(c1) << 16);
  }
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r = lane & 15, q = lane >> 4;
  const size_t row0 = (size_t)blockIdx.x * kMixRows;
  const size_t split = blockIdx.y;
  const size_t k0 = split * red_per_split;
  const size_t k1 = min(k0 + red_per_split, red_len);

  uint4 pa[A_LOADS];
  uint4 pb[B_LOADS];
  auto gload = [&](size_t kk) {
#pragma unroll
    for (int j = 0; j < A_LOADS; ++j) {
      const int e = tid + 256 * j;
      const int row = e >> 1, half = e & 1;
      pa[j] = *reinterpret_cast<const uint4*>(codes + (row0 + row) * code_stride + (kk >> 2) + 16 * half);
    }
#pragma unroll
    for (int j = 0; j < B_LOADS; ++j) {
      const int e = tid + 256 * j;
      if (e < 2 * KP * B_SEGS) {
        const int hl = e / (KP * B_SEGS), rem = e % (KP * B_SEGS);
        const int col = rem / B_SEGS, seg = rem % B_SEGS;
        const __bf16* src = (hl ? blo : bhi) + (size_t)col * b_stride + kk + 8 * seg;
        pb[j] = *reinterpret_cast<const uint4*>(src);
      }
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int j = 0; j < A_LOADS; ++j) {
      const int e = tid + 256 * j;
      const int row = e >> 1, half = e & 1;
      uint2* d = reinterpret_cast<uint2*>(sA + row * kArow + 16 * half);
      d[0] = make_uint2(pa[j].x, pa[j].y);
      d[1] = make_uint2(pa[j].z, pa[j].w);
    }
#pragma unroll
    for (int j = 0; j < B_LOADS; ++j) {
      const int e = tid + 256 * j;
      if (e < 2 * KP * B_SEGS) {
        const int hl = e / (KP * B_SEGS), rem = e % (KP * B_SEGS);
        const int col = rem / B_SEGS, seg = rem % B_SEGS;
        *reinterpret_cast<uint4*>(sB + (hl * KP + col) * kBcol + 8 * seg) = pb[j];
      }
    }
  };

  floatx4 acc[kMixRT][KT];
#pragma unroll
  for (int g = 0; g < kMixRT; ++g)
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[g][t] = floatx4{0.f, 0.f, 0.f, 0.f};

  if (k0 < k1) gload(k0);
  for (size_t kk = k0; kk < k1; kk += kKC) {
    __syncthreads();  // previous chunk fully consumed (and lut ready on the first pass)
    lstore();
    __syncthreads();
    if (kk + kKC < k1) gload(kk + kKC);
#pragma unroll
    for (int ks = 0; ks < kKC / 32; ++ks) {
      bf16x8 bh[KT], bl[KT];
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const int off = (16 * t + r) * kBcol + ks * 32 + 8 * q;
        bh[t] = *reinterpret_cast<const bf16x8*>(sB + off);
        bl[t] = *reinterpret_cast<const bf16x8*>(sB + KP * kBcol + off);
      }
#pragma unroll
      for (int g = 0; g < kMixRT; ++g) {
        const int row = wave * 16 * kMixRT + 16 * g + r;
        const uint32_t bits = *reinterpret_cast<const uint16_t*>(sA + row * kArow + ks * 8 + 2 * q);
        const bf16x8 a = decode8(bits, lut);
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[t], acc[g][t], 0, 0, 0);
          acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[t], acc[g][t], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < kMixRT; ++g)
#pragma unroll
    for (int t = 0; t < KT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        part[(split * rows_pad + row0 + wave * 16 * kMixRT + 16 * g + 4 * q + j) * KP + 16 * t + r] =
            acc[g][t][j];
}

// ---------------------------------------------------------------------------------
// Small fp64 kernels (deterministic fixed-order reductions)
// ---------------------------------------------------------------------------------
// per-column block partial sums of a col-major (ld) matrix over rows [0, len)
__global__ void col_partial_kernel(const double* __restrict__ v, size_t ld, size_t len,
                                   size_t nchunks, double* __restrict__ part) {
  const size_t c = blockIdx.y;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, len);
  double s = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) s += v[i + c * ld];
  __shared__ double sh[kDotBlock];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c * nchunks + blockIdx.x] = sh[0];
}
__global__ void col_finalize_kernel(const double* __restrict__ part, size_t nchunks, size_t k,
                                    double scale, double* __restrict__ out) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double s = 0.0;
  for (size_t q = 0; q < nchunks; ++q) s += part[c * nchunks + q];
  out[c] = s * scale;
}

// centre and split: x' = x - mean_c; hi = bf16(x'), lo = bf16(x' - hi); zero for rows >= len
// or columns >= k.  Also writes x'_split = hi + lo (fp64) for the exact rank-one terms.
__global__ void centre_split_kernel(const double* __restrict__ x, size_t ldx, size_t len,
                                    size_t k, const double* __restrict__ mean, size_t pad_len,
                                    __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                    double* __restrict__ xs, size_t ldxs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (i >= pad_len) return;
  double v = 0.0;
  if (i < len && c < k) v = x[i + c * ldx] - mean[c];
  const __bf16 h = (__bf16)(float)v;
  const double rem = v - (double)(float)h;
  const __bf16 l = (__bf16)(float)rem;
  hi[c * pad_len + i] = h;
  lo[c * pad_len + i] = l;
  if (xs && i < len) xs[i + c * ldxs] = (double)(float)h + (double)(float)l;
}

// Stage-1 epilogue: T_mc = s_m (U'_mc - 2p_m cs_c) + zs_m mean_c; W_mc = s_m T_mc (col-major)
template <int KP>
__global__ void stage1_epilogue_kernel(const float* __restrict__ part, size_t splits, size_t m_pad,
                                       size_t m, size_t k, const double* __restrict__ sc,
                                       const double* __restrict__ p2, const double* __restrict__ zs,
                                       const double* __restrict__ cs, const double* __restrict__ mean,
                                       double* __restrict__ w, size_t ldw) {
  const size_t mm = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (mm >= m || c >= k) return;
  double u = 0.0;
  for (size_t q = 0; q < splits; ++q) u += (double)part[(q * m_pad + mm) * KP + c];
  const double s = sc[mm];
  const double t = s * (u - p2[mm] * cs[c]) + zs[mm] * mean[c];
  w[mm + c * ldw] = s * t;
}

// per-column dot with a fixed weight vector over [0, len): partials
__global__ void wdot_partial_kernel(const double* __restrict__ x, size_t ld,
                                    const double* __restrict__ wt, size_t len, size_t nchunks,
                                    double* __restrict__ part) {
  const size_t c = blockIdx.y;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, len);
  double s = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) s = fma(wt[i], x[i + c * ld], s);
  __shared__ double sh[kDotBlock];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c * nchunks + blockIdx.x] = sh[0];
}

// Stage-2 epilogue: M Y_ic = sum_s part + nu_c (r_i - P2) - S2p_c; optional Sigma form.
template <int KP>
__global__ void stage2_epilogue_kernel(const float* __restrict__ part, size_t splits,
                                       size_t n_pad, size_t n, size_t k,
                                       const double* __restrict__ rsum, double P2,
                                       const double* __restrict__ nu,
                                       const double* __restrict__ s2p, double inv_m,
                                       double* __restrict__ out, size_t ldo,
                                       const double* __restrict__ v, size_t ldv,
                                       const double* __restrict__ wts, double tau0, double tau1) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (i >= n || c >= k) return;
  double y = 0.0;
  for (size_t q = 0; q < splits; ++q) y += (double)part[(q * n_pad + i) * KP + c];
  y = (y + nu[c] * (rsum[i] - P2) - s2p[c]) * inv_m;
  if (wts) y = tau0 * v[i + c * ldv] / fmax(wts[i], 1e-30) + tau1 * y;
  out[i + c * ldo] = y;
}

// ---------------------------------------------------------------------------------
// Missing calls in the factored forms.
// ---------------------------------------------------------------------------------
// A missing call (code 01) decodes to 0 in the MFMA operand, while saige-rs imputes it to
// the mean 2p_m, so that its standardized value is 0 (pcg.rs:108-113).  With per-marker
// lists of the missing samples and per-sample lists of the missing markers, two sparse
// corrections make the factored product exact again:
//   stage 1:  T_mc += s_m 2p_m sum_{i missing in m} (x'_split_ic + mean_c)
//             (so W_mc = s_m T_mc gains s_m^2 2p_m times that sum);
//   stage 2:  M Y_ic += sum_{m missing in i} 2p_m (W'_split_mc + nu_c).
// (r_i and ac_m already count only the calls present.)  Each correction is a gather of
// k-long rows of a row-major copy of the split operand, summed in list order.

// Missing calls among the first `len` codes of each row (rows of rb bytes): count per row.
__global__ __launch_bounds__(256) void miss_count_kernel(const uint8_t* __restrict__ codes,
                                                         size_t rb, size_t len,
                                                         unsigned* __restrict__ cnt) {
  const size_t row = blockIdx.x;
  const size_t nw = (len + 15) / 16;
  unsigned c = 0;
  for (size_t w = threadIdx.x; w < nw; w += 256) {
    const uint32_t word = *reinterpret_cast<const uint32_t*>(codes + row * rb + 4 * w);
    const size_t valid = len - 16 * w;
    const uint32_t mask = valid >= 16 ? 0x55555555u : ((1u << (2 * valid)) - 1u) & 0x55555555u;
    c += __popc(word & ~(word >> 1) & mask);
  }
  __shared__ unsigned sh[256];
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) cnt[row] = sh[0];
}

// The indices of those missing calls, in increasing order: row rows[b] writes from ptr[b].
__global__ __launch_bounds__(256) void miss_fill_kernel(const uint8_t* __restrict__ codes,
                                                        size_t rb, size_t len,
                                                        const uint32_t* __restrict__ rows,
                                                        const uint64_t* __restrict__ ptr,
                                                        uint32_t* __restrict__ out) {
  const size_t row = rows[blockIdx.x];
  const size_t nw = (len + 15) / 16;
  __shared__ unsigned sc[256];
  uint64_t base = ptr[blockIdx.x];
  for (size_t w0 = 0; w0 < nw; w0 += 256) {
    const size_t w = w0 + threadIdx.x;
    uint32_t bits = 0;
    if (w < nw) {
      const uint32_t word = *reinterpret_cast<const uint32_t*>(codes + row * rb + 4 * w);
      const size_t valid = len - 16 * w;
      const uint32_t mask = valid >= 16 ? 0x55555555u : ((1u << (2 * valid)) - 1u) & 0x55555555u;
      bits = word & ~(word >> 1) & mask;
    }
    const unsigned c = __popc(bits);
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {  // inclusive scan
      const unsigned add = (int)threadIdx.x >= s ? sc[threadIdx.x - s] : 0u;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    uint64_t pos = base + sc[threadIdx.x] - c;
    while (bits) {
      const int b = __builtin_ctz(bits);
      out[pos++] = (uint32_t)(16 * w + (size_t)(b >> 1));
      bits &= bits - 1u;
    }
    base += sc[255];
    __syncthreads();
  }
}

// Sparse lists of one side: the rows with missing calls, their offsets and the indices.
struct MissList {
  DevBuf ids, ptr, list;
  size_t rows = 0;
};

static int build_miss_list(hipStream_t s, const uint8_t* codes, size_t rb, size_t rows,
                           size_t len, MissList& ml) {
  DevBuf cnt;
  SAIGE_TRY(cnt.ensure(rows * 4));
  miss_count_kernel<<<(unsigned)rows, 256, 0, s>>>(codes, rb, len, cnt.as<unsigned>());
  SAIGE_HIP(hipGetLastError());
  std::vector<unsigned> hc(rows);
  SAIGE_HIP(hipMemcpyAsync(hc.data(), cnt.p, rows * 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> ids;
  std::vector<uint64_t> ptr(1, 0);
  for (size_t r = 0; r < rows; ++r)
    if (hc[r]) {
      ids.push_back((uint32_t)r);
      ptr.push_back(ptr.back() + hc[r]);
    }
  ml.rows = ids.size();
  if (ml.rows == 0) return SAIGE_OK;
  SAIGE_TRY(ml.ids.ensure(ids.size() * 4));
  SAIGE_TRY(ml.ptr.ensure(ptr.size() * 8));
  SAIGE_TRY(ml.list.ensure(ptr.back() * 4));
  SAIGE_HIP(hipMemcpyAsync(ml.ids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(ml.ptr.p, ptr.data(), ptr.size() * 8, hipMemcpyHostToDevice, s));
  miss_fill_kernel<<<(unsigned)ml.rows, 256, 0, s>>>(codes, rb, len, ml.ids.as<uint32_t>(),
                                                     ml.ptr.as<uint64_t>(), ml.list.as<uint32_t>());
  SAIGE_HIP(hipGetLastError());
  SAIGE_HIP(hipStreamSynchronize(s));  // the host vectors end here
  return SAIGE_OK;
}

// means[c] = (sum of the column's chunk partials) * inv_len
__global__ void chunk_means_kernel(const double* __restrict__ psum, size_t nch, size_t k,
                                   double inv_len, double* __restrict__ means) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double s = 0.0;
  for (size_t q = 0; q < nch; ++q) s += psum[c * nch + q];
  means[c] = s * inv_len;
}

// rows[i * k + c] = xs[i + c * ldxs] + mean[c]: the split operand as rows, for the gathers
__global__ void rows_plus_mean_kernel(const double* __restrict__ xs, size_t ldxs, size_t len,
                                      size_t k, const double* __restrict__ mean,
                                      double* __restrict__ rows) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len * k) return;
  const size_t i = e / k, c = e % k;
  rows[e] = xs[i + c * ldxs] + mean[c];
}

// w[m + c * ldw] += sc_m^2 2p_m sum_{i in list(m)} rows[i * k + c] for the listed markers
__global__ void miss_stage1_kernel(const uint32_t* __restrict__ ids,
                                   const uint64_t* __restrict__ ptr,
                                   const uint32_t* __restrict__ list,
                                   const double* __restrict__ rows, size_t k,
                                   const double* __restrict__ sc, const double* __restrict__ p2,
                                   double* __restrict__ w, size_t ldw) {
  const size_t mk = ids[blockIdx.x];
  const uint64_t e0 = ptr[blockIdx.x], e1 = ptr[blockIdx.x + 1];
  const double f = sc[mk] * sc[mk] * p2[mk];
  for (size_t c = threadIdx.x; c < k; c += blockDim.x) {
    double acc = 0.0;
    for (uint64_t e = e0; e < e1; ++e) acc += rows[(size_t)list[e] * k + c];
    w[mk + c * ldw] += f * acc;
  }
}

// out[i + c * ldo] += scale * sum_{m in list(i)} 2p_m rows[m * k + c] for the listed samples
__global__ void miss_stage2_kernel(const uint32_t* __restrict__ ids,
                                   const uint64_t* __restrict__ ptr,
                                   const uint32_t* __restrict__ list,
                                   const double* __restrict__ rows, size_t k,
                                   const double* __restrict__ p2, double scale,
                                   double* __restrict__ out, size_t ldo) {
  const size_t i = ids[blockIdx.x];
  const uint64_t e0 = ptr[blockIdx.x], e1 = ptr[blockIdx.x + 1];
  for (size_t c = threadIdx.x; c < k; c += blockDim.x) {
    double acc = 0.0;
    for (uint64_t e = e0; e < e1; ++e) {
      const size_t mk = list[e];
      acc += p2[mk] * rows[mk * k + c];
    }
    out[i + c * ldo] += scale * acc;
  }
}

// ---------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------
struct MixedState {
  DevBuf geno_t, rsum, sc, p2, zs, nmiss;
  double P2 = 0.0;
  size_t rbt = 0;
  // per-product workspaces
  DevBuf vhi, vlo, vs, whi, wlo, ws_, w, part1, part2, red, cs, mean, nu, s2p;
  // int8 path: sliced operands, per-column slice scales, integer partials
  DevBuf vq, wq, ipart1, ipart2, cols;
  // missing calls: per-marker sample lists, per-sample marker lists, gather rows
  MissList miss_mk, miss_sm;
  DevBuf mrows, mmean;
};

static MixedState* mstate(saige_grm* h);
MixedState* mixed_state(saige_grm* h) { return mstate(h); }

// Stage-1 missing-call correction of W (m x k, ldw) from the split V' (xs: n x k, ldxs) and
// the column means of V (`mean`, or the chunk partials psum over nch chunks).
static int missing_stage1(saige_grm* h, const double* xs, size_t ldxs, size_t k,
                          const double* mean, const double* psum, size_t nch, double* w,
                          size_t ldw) {
  MixedState* ms = mstate(h);
  if (ms->miss_mk.rows == 0 || k == 0) return SAIGE_OK;
  hipStream_t s = h->stream;
  const size_t n = h->g.n;
  SAIGE_TRY(ms->mrows.ensure(n * k * 8));
  if (!mean) {
    SAIGE_TRY(ms->mmean.ensure(k * 8));
    chunk_means_kernel<<<(unsigned)ceil_div(k, 256), 256, 0, s>>>(psum, nch, k, 1.0 / (double)n,
                                                                   ms->mmean.as<double>());
    mean = ms->mmean.as<double>();
  }
  rows_plus_mean_kernel<<<(unsigned)ceil_div(n * k, 256), 256, 0, s>>>(xs, ldxs, n, k, mean,
                                                                        ms->mrows.as<double>());
  miss_stage1_kernel<<<(unsigned)ms->miss_mk.rows, 64, 0, s>>>(
      ms->miss_mk.ids.as<uint32_t>(), ms->miss_mk.ptr.as<uint64_t>(),
      ms->miss_mk.list.as<uint32_t>(), ms->mrows.as<double>(), k, ms->sc.as<double>(),
      ms->p2.as<double>(), w, ldw);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

// Stage-2 missing-call correction of the output (scale = tau1 / M, or 1 / M) from the split
// W' (ws: m x k, ldws) and the column means of W (`nu`, or the chunk partials psum).
static int missing_stage2(saige_grm* h, const double* ws, size_t ldws, size_t k,
                          const double* nu, const double* psum, size_t nch, double* out,
                          size_t ldo, double scale) {
  MixedState* ms = mstate(h);
  if (ms->miss_sm.rows == 0 || k == 0) return SAIGE_OK;
  hipStream_t s = h->stream;
  const size_t m = h->g.m;
  SAIGE_TRY(ms->mrows.ensure(m * k * 8));
  if (!nu) {
    SAIGE_TRY(ms->mmean.ensure(k * 8));
    chunk_means_kernel<<<(unsigned)ceil_div(k, 256), 256, 0, s>>>(psum, nch, k, 1.0 / (double)m,
                                                                   ms->mmean.as<double>());
    nu = ms->mmean.as<double>();
  }
  rows_plus_mean_kernel<<<(unsigned)ceil_div(m * k, 256), 256, 0, s>>>(ws, ldws, m, k, nu,
                                                                        ms->mrows.as<double>());
  miss_stage2_kernel<<<(unsigned)ms->miss_sm.rows, 64, 0, s>>>(
      ms->miss_sm.ids.as<uint32_t>(), ms->miss_sm.ptr.as<uint64_t>(),
      ms->miss_sm.list.as<uint32_t>(), ms->mrows.as<double>(), k, ms->p2.as<double>(), scale,
      out, ldo);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

int grm_mixed_prepare(saige_grm* h) {
  GrmDevice& g = h->g;
  hipStream_t s = h->stream;
  MixedState* ms = mstate(h);
  ms->rbt = g.m_pad / 4;
  SAIGE_TRY(ms->geno_t.ensure(g.n_pad * ms->rbt));
  SAIGE_TRY(ms->rsum.ensure(g.n_pad * 8));
  SAIGE_TRY(ms->nmiss.ensure(64 + g.n_pad * 4));
  unsigned* d_rsum_u = reinterpret_cast<unsigned*>(ms->nmiss.as<char>() + 64);
  SAIGE_HIP(hipMemsetAsync(ms->nmiss.p, 0, 64 + g.n_pad * 4, s));
  {
    // about 2,048 work-groups: sample blocks x marker groups
    const size_t tiles = g.m_pad / 256, sblocks = g.n_pad / 256;
    const size_t tpg = ceil_div(tiles, std::max<size_t>(1, std::min(tiles, ceil_div(2048, sblocks))));
    transpose_codes_kernel<<<dim3((unsigned)ceil_div(tiles, tpg), (unsigned)sblocks), 256, 0, s>>>(
        g.geno, g.row_bytes, ms->geno_t.as<uint8_t>(), ms->rbt, g.n, g.m, g.m_pad, tpg, d_rsum_u,
        ms->nmiss.as<unsigned>());
  }
  u32_to_f64_kernel<<<(unsigned)ceil_div(g.n_pad, 256), 256, 0, s>>>(d_rsum_u, g.n_pad,
                                                                     ms->rsum.as<double>());
  SAIGE_HIP(hipGetLastError());
  unsigned nmiss = 0;
  SAIGE_HIP(hipMemcpyAsync(&nmiss, ms->nmiss.p, 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  if (nmiss) {
    // per-marker lists from the SNP-major rows (real samples), per-sample lists from the
    // sample-major copy (real markers); both before the int8 forms re-block the code matrices
    SAIGE_TRY(build_miss_list(s, g.geno, g.row_bytes, g.m, g.n, ms->miss_mk));
    SAIGE_TRY(build_miss_list(s, ms->geno_t.as<uint8_t>(), ms->rbt, g.n, g.m, ms->miss_sm));
  }
  // per-marker s_m, 2 p_m (from the host's 1/sd and 2p/sd) and, on the device,
  // zs_m = s_m (ac_m - 2 p_m n) with ac from the stored codes
  std::vector<double> sc(g.m_pad, 0.0), p2(g.m_pad, 0.0);
  double P2 = 0.0;
  for (size_t j = 0; j < g.m; ++j) {
    sc[j] = h->host_scale[j];
    p2[j] = sc[j] > 0.0 ? h->host_shift[j] / sc[j] : 0.0;  // shift = 2p * s
    P2 += p2[j];
  }
  ms->P2 = P2;
  SAIGE_TRY(ms->sc.ensure(g.m_pad * 8));
  SAIGE_TRY(ms->p2.ensure(g.m_pad * 8));
  SAIGE_TRY(ms->zs.ensure(g.m_pad * 8));
  SAIGE_HIP(hipMemcpyAsync(ms->sc.p, sc.data(), g.m_pad * 8, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(ms->p2.p, p2.data(), g.m_pad * 8, hipMemcpyHostToDevice, s));
  {
    DevBuf d_ac;
    SAIGE_TRY(d_ac.ensure(std::max<size_t>(g.m, 1) * 8));
    SAIGE_TRY(marker_ac_from_device(s, g.geno, g.row_bytes, g.n, g.m, d_ac.as<double>()));
    marker_zs_kernel<<<(unsigned)ceil_div(g.m_pad, 256), 256, 0, s>>>(
        ms->sc.as<double>(), ms->p2.as<double>(), d_ac.as<double>(), (double)g.n, g.m, g.m_pad,
        ms->zs.as<double>());
    SAIGE_HIP(hipGetLastError());
    SAIGE_HIP(hipStreamSynchronize(s));  // the host vectors and d_ac end here
  }
  if (h->precision == SAIGE_PREC_INT8 || h->precision == SAIGE_PREC_INT8X2) {
    // nothing reads the row-major forms from here on
    SAIGE_TRY(block_codes(s, h->geno, g.m_pad, g.row_bytes));
    g.geno = h->geno.as<uint8_t>();
    SAIGE_TRY(block_codes(s, ms->geno_t, g.n_pad, ms->rbt));
  }
  return SAIGE_OK;
}

template <int KT>
static int mixed_product_kt(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                            size_t ldo, const double* d_w, const double* tau, size_t c0) {
  constexpr int KP = 16 * KT;
  GrmDevice& g = h->g;
  MixedState* ms = mstate(h);
  hipStream_t s = h->stream;
  const size_t n = g.n, m = g.m, n_pad = g.n_pad, m_pad = g.m_pad;
  const size_t nch_n = std::max<size_t>(1, ceil_div(n, kDotChunk));
  const size_t nch_m = std::max<size_t>(1, ceil_div(m, kDotChunk));
  SAIGE_TRY(ms->vhi.ensure(KP * n_pad * 2));
  SAIGE_TRY(ms->vlo.ensure(KP * n_pad * 2));
  SAIGE_TRY(ms->vs.ensure(n * KP * 8));
  SAIGE_TRY(ms->whi.ensure(KP * m_pad * 2));
  SAIGE_TRY(ms->wlo.ensure(KP * m_pad * 2));
  SAIGE_TRY(ms->ws_.ensure(m * KP * 8));
  SAIGE_TRY(ms->w.ensure(m * KP * 8));
  SAIGE_TRY(ms->red.ensure(std::max(nch_n, nch_m) * KP * 8));
  SAIGE_TRY(ms->cs.ensure(KP * 8));
  SAIGE_TRY(ms->mean.ensure(KP * 8));
  SAIGE_TRY(ms->nu.ensure(KP * 8));
  SAIGE_TRY(ms->s2p.ensure(KP * 8));
  double* red = ms->red.as<double>();
  double* mean = ms->mean.as<double>();
  double* cs = ms->cs.as<double>();
  double* nu = ms->nu.as<double>();
  double* s2p = ms->s2p.as<double>();
  const unsigned kb = 1;

  // ---- V: column means, centre + split, cs' = colsum of the split ----
  col_partial_kernel<<<dim3((unsigned)nch_n, (unsigned)k), kDotBlock, 0, s>>>(d_v, ldv, n, nch_n, red);
  col_finalize_kernel<<<kb, 256, 0, s>>>(red, nch_n, k, 1.0 / (double)n, mean);
  centre_split_kernel<<<dim3((unsigned)ceil_div(n_pad, 256), (unsigned)KP), 256, 0, s>>>(
      d_v, ldv, n, k, mean, n_pad, ms->vhi.as<__bf16>(), ms->vlo.as<__bf16>(), ms->vs.as<double>(), n);
  col_partial_kernel<<<dim3((unsigned)nch_n, (unsigned)k), kDotBlock, 0, s>>>(ms->vs.as<double>(), n,
                                                                               n, nch_n, red);
  col_finalize_kernel<<<kb, 256, 0, s>>>(red, nch_n, k, 1.0, cs);

  // ---- stage 1: U' = G^T V'_split (rows = markers, reduction over samples) ----
  const size_t rblk1 = m_pad / kMixRows;
  size_t splits1 = std::max<size_t>(1, std::min(ceil_div(1024, rblk1), n_pad / 512));
  size_t rps1 = round_up(ceil_div(n_pad, splits1), kKC);
  splits1 = ceil_div(n_pad, rps1);
  SAIGE_TRY(ms->part1.ensure(splits1 * m_pad * KP * 4));
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)KP, (uint32_t)k}, mix_gemm_lds_kernel<KT>, dim3((unsigned)rblk1, (unsigned)splits1),
                         dim3(256), s, g.geno, g.row_bytes, ms->vhi.as<__bf16>(),
                         ms->vlo.as<__bf16>(), n_pad, rps1, n_pad, ms->part1.as<float>(), m_pad));
  stage1_epilogue_kernel<KP><<<dim3((unsigned)ceil_div(m, 256), (unsigned)k), 256, 0, s>>>(
      ms->part1.as<float>(), splits1, m_pad, m, k, ms->sc.as<double>(), ms->p2.as<double>(),
      ms->zs.as<double>(), cs, mean, ms->w.as<double>(), m);
  SAIGE_TRY(missing_stage1(h, ms->vs.as<double>(), n, k, mean, nullptr, 0, ms->w.as<double>(), m));
  LocoCodes lc;
  if (loco_codes(h, c0, k, lc)) loco_mask(h, ms->w.as<double>(), 1, m, m, k, lc, s);

  // ---- W: mean over real markers, centre + split, S2p = sum_m 2p_m W'_split ----
  col_partial_kernel<<<dim3((unsigned)nch_m, (unsigned)k), kDotBlock, 0, s>>>(ms->w.as<double>(), m,
                                                                               m, nch_m, red);
  col_finalize_kernel<<<kb, 256, 0, s>>>(red, nch_m, k, 1.0 / (double)m, nu);
  centre_split_kernel<<<dim3((unsigned)ceil_div(m_pad, 256), (unsigned)KP), 256, 0, s>>>(
      ms->w.as<double>(), m, m, k, nu, m_pad, ms->whi.as<__bf16>(), ms->wlo.as<__bf16>(),
      ms->ws_.as<double>(), m);
  wdot_partial_kernel<<<dim3((unsigned)nch_m, (unsigned)k), kDotBlock, 0, s>>>(
      ms->ws_.as<double>(), m, ms->p2.as<double>(), m, nch_m, red);
  col_finalize_kernel<<<kb, 256, 0, s>>>(red, nch_m, k, 1.0, s2p);

  // ---- stage 2: G W'_split (rows = samples, reduction over markers) ----
  const size_t rblk2 = n_pad / kMixRows;
  size_t splits2 = std::max<size_t>(1, std::min(ceil_div(1024, rblk2), m_pad / 256));
  size_t rps2 = round_up(ceil_div(m_pad, splits2), kKC);
  splits2 = ceil_div(m_pad, rps2);
  SAIGE_TRY(ms->part2.ensure(splits2 * n_pad * KP * 4));
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)KP, (uint32_t)k}, mix_gemm_lds_kernel<KT>, dim3((unsigned)rblk2, (unsigned)splits2),
                         dim3(256), s, ms->geno_t.as<uint8_t>(), ms->rbt, ms->whi.as<__bf16>(),
                         ms->wlo.as<__bf16>(), m_pad, rps2, m_pad, ms->part2.as<float>(), n_pad));
  stage2_epilogue_kernel<KP><<<dim3((unsigned)ceil_div(n, 256), (unsigned)k), 256, 0, s>>>(
      ms->part2.as<float>(), splits2, n_pad, n, k, ms->rsum.as<double>(), ms->P2, nu, s2p,
      1.0 / (double)m, d_out, ldo, d_v, ldv, d_w, tau ? tau[0] : 0.0, tau ? tau[1] : 0.0);
  SAIGE_HIP(hipGetLastError());
  SAIGE_TRY(missing_stage2(h, ms->ws_.as<double>(), m, k, nu, nullptr, 0, d_out, ldo,
                           (tau ? tau[1] : 1.0) / (double)m));
  h->products += k;
  return SAIGE_OK;
}

int grm_int8_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                     size_t ldo, const double* d_w, const double* tau);

int grm_mixed_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                      size_t ldo, const double* d_w, const double* tau) {
  if (h->precision == SAIGE_PREC_INT8 || h->precision == SAIGE_PREC_INT8X2)
    return grm_int8_product(h, d_v, ldv, k, d_out, ldo, d_w, tau);
  for (size_t c0 = 0; c0 < k; c0 += 128) {
    const size_t kc = std::min<size_t>(128, k - c0);
    const double* v = d_v + c0 * ldv;
    double* o = d_out + c0 * ldo;
    int rc;
    if (kc <= 16) rc = mixed_product_kt<1>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 32) rc = mixed_product_kt<2>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 48) rc = mixed_product_kt<3>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 64) rc = mixed_product_kt<4>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else if (kc <= 96) rc = mixed_product_kt<6>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    else rc = mixed_product_kt<8>(h, v, ldv, kc, o, ldo, d_w, tau, c0);
    if (rc != SAIGE_OK) return rc;
  }
  return SAIGE_OK;
}

// one MixedState per handle, owned through the handle's geno_t buffer slot
static MixedState* mstate(saige_grm* h) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (!h->mixed) h->mixed = new MixedState();
  return static_cast<MixedState*>(h->mixed);
}

void grm_mixed_release(saige_grm* h) {
  delete static_cast<MixedState*>(h->mixed);
  h->mixed = nullptr;
}


// =================================================================================
// Int8 paths (SAIGE_PREC_INT8, SAIGE_PREC_INT8X2): exact genotypes times int8 operand
// slices on integer MFMA.
// =================================================================================
// The centred operand x' of each column is split into 2 or 3 int8 slices with per-column
// scales s0 = max|x'| / 127, s1 = s0 / 254, s2 = s1 / 254:
//   q0 = rint(x'/s0),  q1 = rint((x' - s0 q0)/s1),  q2 = rint((x' - s0 q0 - s1 q1)/s2),
// each in [-127, 127], so |x' - sum_j s_j q_j| <= s_{S-1}/2: max|x'| / (2 * 127 * 254^2),
// about 6.1e-8 of the column's largest entry, with three slices; max|x'| / (2 * 127 * 254),
// about 1.55e-5, with two.  SAIGE_PREC_INT8 handles always take three slices;
// SAIGE_PREC_INT8X2 handles take two unless a product asks for three on some of its
// columns (saige_grm::slice_req).  G^T Q_j is computed exactly with
// v_mfma_i32_16x16x64_i8 (int32 accumulation: |sum| <= 2 * 127 * n < 2^31 for n <= 8e6),
// the slices are recombined in fp64, and x'_split = sum_j s_j q_j enters the rank-one
// terms of the factored form exactly as in the bf16 path.  Integer partial sums make the
// result independent of how the reduction is split, and a column's result depends only on
// its own slices, not on the other columns of the launch.
//
// Operand layout of one launch with k vectors, k3 of them with a third slice: slice 0 of
// vector c is column c, slice 1 column k + c, slice 2 column 2k + (the rank of c among the
// three-slice vectors); columns from 2k + k3 up to the padded width KPT are zero.  Each
// column is pad_len long in reduction order.  Within every group of 16 reduction indices the
// bytes are stored 4x4-transposed (index 4b+u at byte 4u+b): a lane expands one 32-bit word
// of packed genotype values (value 4b+u at bits 8b+2u) by a shift and a mask into four
// words, word u holding values {u, 4+u, 8+u, 12+u} one per byte, and the operand's bytes
// then pair with the same reduction indices.
//
// Column groups: KPT = G * 16 * KT with KT <= 10 tiles per group.  One launch covers all G
// groups; work-groups that share a row block and a reduction split but differ in the group
// are issued back to back on the same XCD (XCD = linear work-group id mod 8), so the
// genotype bytes that one of them streams from HBM are read by the others from that XCD's
// L2: the genotypes stream once per pass, whatever the number of columns.
constexpr int kI8MaxVec = 256;    // vectors per launch
constexpr int kI8MaxCols = 512;   // slice columns per launch
constexpr size_t kI8MaxRed = 8000000;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Bit c set: vector c of the launch has a third operand slice.
struct SliceMap {
  uint32_t bits[kI8MaxVec / 32];
};
// Operand column of the third slice of vector c, or -1 without one.
__host__ __device__ __forceinline__ int third_col(const SliceMap& sm, size_t k, size_t c) {
  const uint32_t w = sm.bits[c >> 5], b = 1u << (c & 31);
  if (!(w & b)) return -1;
  int r = __builtin_popcount(w & (b - 1));
  for (size_t j = 0; j < (c >> 5); ++j) r += __builtin_popcount(sm.bits[j]);
  return (int)(2 * k) + r;
}

// Column scalars rebuilt from fixed-order partials (one value per kDotChunk rows, combined in
// chunk order): the mean of column c, and the three slice scales from the largest
// |x - mean|, which is max(xmax - mean, mean - xmin) over the chunks' maxima and minima
// (rounding is monotone and odd, so this is the largest rounded |x_i - mean| itself);
// s0 = max / 127, s1 = s0 / 254, s2 = s1 / 254, and 1 for an all-zero column.
__device__ __forceinline__ double chunk_sum(const double* __restrict__ part, size_t c,
                                            size_t nchunks) {
  double s = 0.0;
  for (size_t q = 0; q < nchunks; ++q) s += part[c * nchunks + q];
  return s;
}
__device__ __forceinline__ void slice_scales(const double* __restrict__ pmax,
                                             const double* __restrict__ pmin, double mean,
                                             size_t c, size_t nchunks, double* sc3) {
  double m = 0.0;
  for (size_t q = 0; q < nchunks; ++q)
    m = fmax(m, fmax(pmax[c * nchunks + q] - mean, mean - pmin[c * nchunks + q]));
  const double s0 = m > 0.0 ? m / 127.0 : 1.0;
  sc3[0] = s0;
  sc3[1] = s0 / 254.0;
  sc3[2] = sc3[1] / 254.0;
}

// Per column and chunk of kDotChunk rows: the sum (the same order as col_partial_kernel),
// the maximum and the minimum.
__global__ __launch_bounds__(kDotBlock) void col_stats_kernel(
    const double* __restrict__ x, size_t ld, size_t len, size_t nchunks,
    double* __restrict__ psum, double* __restrict__ pmax, double* __restrict__ pmin) {
  const size_t c = blockIdx.y;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, len);
  double s = 0.0, hi = -INFINITY, lo = INFINITY;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    const double v = x[i + c * ld];
    s += v;
    hi = fmax(hi, v);
    lo = fmin(lo, v);
  }
  __shared__ double sh[3][kDotBlock];
  sh[0][threadIdx.x] = s;
  sh[1][threadIdx.x] = hi;
  sh[2][threadIdx.x] = lo;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
      sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum[c * nchunks + blockIdx.x] = sh[0][0];
    pmax[c * nchunks + blockIdx.x] = sh[1][0];
    pmin[c * nchunks + blockIdx.x] = sh[2][0];
  }
}

__device__ __forceinline__ size_t perm16(size_t i) {
  return (i & ~(size_t)15) | ((i & 3) << 2) | ((i >> 2) & 3);
}
__device__ __forceinline__ double q127(double t) { return fmin(fmax(rint(t), -127.0), 127.0); }

// x' = x - mean -> 2 or 3 int8 slices per column (stored in the permuted reduction order,
// zero for rows >= len), the fp64 value they represent (xs), and per chunk of kDotChunk rows
// the partial sum of wt * xs (wt = 1 when null) for the rank-one terms.  One block per
// (chunk, column); grid.y = k + (KPT - 2k - k3): the last rows of the grid clear the padding
// columns.
// The column's mean and slice scales, finished once per operand pass from the fixed-order
// partials: ms4[4c .. 4c + 3] = mean, s0, s1, s2.
__global__ void col_mean_scales_kernel(size_t k, const double* __restrict__ psum,
                                       const double* __restrict__ pmax,
                                       const double* __restrict__ pmin, double inv_len,
                                       size_t nchunks, double* __restrict__ ms4) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  const double mean = chunk_sum(psum, c, nchunks) * inv_len;
  ms4[4 * c] = mean;
  slice_scales(pmax, pmin, mean, c, nchunks, ms4 + 4 * c + 1);
}

__global__ __launch_bounds__(kDotBlock) void quantize_slices_kernel(
    const double* __restrict__ x, size_t ldx, size_t len, size_t k, const SliceMap sm,
    size_t k3, const double* __restrict__ ms4, size_t nchunks, size_t pad_len,
    int8_t* __restrict__ q, double* __restrict__ xs, size_t ldxs,
    const double* __restrict__ wt, double* __restrict__ pdot) {
  const size_t c = blockIdx.y;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, pad_len);
  if (c >= k) {  // padding column 2k + k3 + (c - k)
    const size_t pc = 2 * k + k3 + (c - k);
    for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) q[pc * pad_len + perm16(i)] = 0;
    return;
  }
  const int c3 = third_col(sm, k, c);
  const double mu = ms4[4 * c], s0 = ms4[4 * c + 1], s1 = ms4[4 * c + 2], s2 = ms4[4 * c + 3];
  (void)nchunks;
  double acc = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (i < len) {
      const double xv = x[i + c * ldx] - mu;
      t0 = q127(xv / s0);
      double r = fma(-s0, t0, xv);
      t1 = q127(r / s1);
      double v;
      if (c3 >= 0) {
        r = fma(-s1, t1, r);
        t2 = q127(r / s2);
        v = s0 * t0 + s1 * t1 + s2 * t2;
      } else {
        v = s0 * t0 + s1 * t1;
      }
      xs[i + c * ldxs] = v;
      acc = wt ? fma(wt[i], v, acc) : acc + v;
    }
    const size_t pos = perm16(i);
    q[c * pad_len + pos] = (int8_t)t0;
    q[(k + c) * pad_len + pos] = (int8_t)t1;
    if (c3 >= 0) q[(size_t)c3 * pad_len + pos] = (int8_t)t2;
  }
  __shared__ double sh[kDotBlock];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && i0 < len) pdot[c * nchunks + blockIdx.x] = sh[0];
}

// 16 bytes from global memory into LDS without passing through registers: lane i's 16 bytes
// land at LDS byte address lds + 16 i (lds is wave-uniform and goes through M0).  The load
// counts in vmcnt like any vector load; its data are in LDS once vmcnt has dropped past it.
// Written in asm so that the compiler neither adds waits of its own between these loads and
// the LDS reads (the ring below orders them with explicit vmcnt waits and barriers) nor
// expects M0 to survive.
__device__ __forceinline__ void lds_dma16(uint32_t lds, const void* g) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds), "v"(g)
               : "memory", "m0");
}
__device__ __forceinline__ uint32_t lds_address(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// LDS stages of the int8 GEMM ring: as many 16 KB + KT * 2 KB stages as fit in 160 KB, at
// most five; four for KT <= 2 (two work-groups per CU for the genotype-bound narrow passes).
template <int KT>
struct I8Ring {
  static constexpr int kStage = (int)kI8Block + 16 * KT * (int)kI8Chunk;
  static constexpr int kFit = 163840 / kStage;
  static constexpr int NS = KT <= 2 ? 4 : (kFit < 5 ? kFit : 5);
};

// part[split][col][row] (split stride part_ld * rows_pad; columns grp*16*KT .. of the group)
// = sum over the split's reduction range of g(row, red) * q[col][red]: column-major per split,
// so that the epilogue reads each column's partials as contiguous rows.
// Work-group order: linear id L runs over round_up(nrb * nsp, 8) * ngroups work-groups;
// j = L / 8 and x = L % 8 give the tile (j / ngroups) * 8 + x (row block tile % nrb, split
// tile / nrb) and the column group j % ngroups, so the groups of one tile are consecutive
// work-groups of one XCD and share its L2 for the genotype blocks.  Tiles past nrb * nsp
// exit at once (a whole work-group).
// Eight waves; wave w owns rows 64 w .. 64 w + 63 of the 512-row block and all 16 KT
// columns of the group (KT <= 10: 160 int32 accumulators per lane at two waves per SIMD).
// Per chunk of 128 reduction indices a stage of the LDS ring holds the 16 KB genotype block
// (layout of block_codes_kernel) and the group's operand chunk, 16 KT columns x 128 bytes,
// its 16-byte units XOR-swizzled by column (unit u of column c at c * 128 + 16 (u ^ (c & 7)))
// so that the 16-byte fragment reads of 8 consecutive columns fall in distinct banks.  Both
// arrive by 16-byte DMA (D of them per wave and chunk; waves past the operand's last unit
// repeat one, which rewrites identical bytes).  The ring runs NS - 2 chunks ahead: chunk
// c + NS - 2 is loaded, spread over chunk c's MFMA stream, into the stage that chunk c - 2
// used, which every wave has left by the barrier that opens chunk c - 1.  The barrier that
// opens chunk c follows each wave's wait for its own loads of chunk c, so after it every
// wave's loads of that chunk are in LDS.  Loads past the range repeat the last chunk into
// stages no wave reads again; all loads are drained before the accumulators are stored.
template <int KT, int NS>
__global__ __launch_bounds__(512) void i8_gemm_kernel(
    const uint8_t* __restrict__ codes, size_t nkc, const int8_t* __restrict__ bq,
    size_t b_stride, size_t cps, int* __restrict__ part, size_t rows_pad, size_t part_ld,
    size_t nrb, size_t nsp, size_t ngroups) {
  constexpr int KCH = (int)kI8Chunk, NW = 8, ROWS = 512, KP = 16 * KT, UPC = KCH / 16;
  constexpr int KS = KCH / 64;                    // MFMA steps per chunk
  constexpr int ABYTES = (int)kI8Block, A_INS = ABYTES / 1024 / NW;
  constexpr int BTOT = KP * KCH / 1024, B_INS = (BTOT + NW - 1) / NW;
  constexpr int STAGE = ABYTES + KP * KCH;
  constexpr int D = A_INS + B_INS;                // DMA loads per wave and chunk
  constexpr int GB = 4 * KS;                      // row-tile blocks of KT MFMAs per chunk
  constexpr int EVERY = GB / D > 0 ? GB / D : 1;  // one DMA load every EVERY blocks
  static_assert(NS >= 3 && NS * STAGE <= 163840 && D * (NS - 3) <= 63, "LDS ring");
  static_assert(ABYTES % (1024 * NW) == 0 && (KP * KCH) % 1024 == 0, "whole DMA loads");
  __shared__ __attribute__((aligned(1024))) int8_t smem[NS * STAGE];
  const size_t L = blockIdx.x, jx = L >> 3;
  const size_t tile = (jx / ngroups) * 8 + (L & 7), grp = jx % ngroups;
  if (tile >= nrb * nsp) return;  // uniform over the work-group
  const int tid = threadIdx.x, lane = tid & 63, r = lane & 15, q = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const size_t rb = tile % nrb, split = tile / nrb;
  const size_t c0 = split * cps, c1 = min(c0 + cps, nkc);
  if (c0 >= c1) return;  // uniform over the work-group
  const size_t nc = c1 - c0;
  const uint8_t* asrc = codes + (rb * nkc + c0) * ABYTES + lane * 16;
  const int8_t* bsrc[B_INS];
  uint32_t bdst[B_INS];
#pragma unroll
  for (int i = 0; i < B_INS; ++i) {
    int j = wave + NW * i;
    if (j >= BTOT) j %= BTOT;
    const int p = 64 * j + lane, c = p / UPC, s = p % UPC, u = s ^ (c & 7);
    bsrc[i] = bq + (grp * KP + c) * b_stride + c0 * KCH + 16 * u;
    bdst[i] = ABYTES + 1024 * j;
  }
  const uint32_t sbase = lds_address(smem);
  auto load = [&](int d, size_t ci, uint32_t st) {
    if (d < A_INS) {
      const int j = wave + NW * d;
      lds_dma16(st + j * 1024, asrc + ci * ABYTES + j * 1024);
    } else {
      lds_dma16(st + bdst[d - A_INS], bsrc[d - A_INS] + ci * KCH);
    }
  };
#pragma unroll
  for (int s = 0; s < NS - 2; ++s)
#pragma unroll
    for (int d = 0; d < D; ++d) load(d, min((size_t)s, nc - 1), sbase + s * STAGE);
  i32x4 acc[4][KT];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[g][t] = i32x4{0, 0, 0, 0};
  int cur = 0, nxt = NS - 2;
  for (size_t c = 0; c < nc; ++c) {
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(D * (NS - 3))
                 : "memory");
    const size_t cn = min(c + NS - 2, nc - 1);
    const uint32_t stn = sbase + nxt * STAGE;
    const int8_t* sa = smem + cur * STAGE + wave * 1024 + lane * 16;
    const int8_t* sb = smem + cur * STAGE + ABYTES + r * KCH;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const u32x4 aw = *reinterpret_cast<const u32x4*>(sa + ks * 8192);
      i32x4 b[KT];
#pragma unroll
      for (int t = 0; t < KT; ++t)
        b[t] = *reinterpret_cast<const i32x4*>(sb + 16 * t * KCH + (((4 * ks + q) ^ (r & 7)) * 16));
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int gb = ks * 4 + g;
        if (gb % EVERY == 0 && gb / EVERY < D) {
          __builtin_amdgcn_sched_barrier(0);
          load(gb / EVERY, cn, stn);
          __builtin_amdgcn_sched_barrier(0);
        }
        // word u holds fields u of the four bytes: reduction indices 4b + u, b = 0..3
        const uint32_t w = aw[g];
        i32x4 a;
        a[0] = (int)(w & 0x03030303u);
        a[1] = (int)((w >> 2) & 0x03030303u);
        a[2] = (int)((w >> 4) & 0x03030303u);
        a[3] = (int)((w >> 6) & 0x03030303u);
#pragma unroll
        for (int t = 0; t < KT; ++t)
          acc[g][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[t], acc[g][t], 0, 0, 0);
      }
    }
    if (GB / EVERY < D) {
#pragma unroll
      for (int d = GB / EVERY; d < D; ++d) load(d, cn, stn);
    }
    cur = cur + 1 == NS ? 0 : cur + 1;
    nxt = nxt + 1 == NS ? 0 : nxt + 1;
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  // lane (r, q) holds rows 16g + 4q .. + 3 of column 16t + r: one 16-byte store each
  const size_t rowb = rb * ROWS + (size_t)wave * 64;
  int* pbase = part + (split * part_ld + grp * KP + r) * rows_pad + rowb + 4 * q;
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int t = 0; t < KT; ++t)
      *reinterpret_cast<i32x4*>(pbase + (size_t)16 * t * rows_pad + 16 * g) = acc[g][t];
}

// Column scalars of the epilogues: [mean or nu, s0, s1, s2, cs or S2p] and the operand
// column of the third slice (-1 without one), finished once per product from the operand
// pass's fixed-order partials (i8_col_scalars_kernel) and read by every epilogue block.
struct I8Col {
  double v[5];
  int c3;
};
__global__ void i8_col_scalars_kernel(size_t k, const SliceMap sm, const double* __restrict__ ms4,
                                      const double* __restrict__ pdot, size_t nch,
                                      I8Col* __restrict__ col) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  I8Col x;
  for (int j = 0; j < 4; ++j) x.v[j] = ms4[4 * c + j];
  x.v[4] = chunk_sum(pdot, c, nch);
  x.c3 = third_col(sm, k, c);
  col[c] = x;
}
// Vectors per epilogue work-group (blockIdx.y): a thread keeps its row's per-row values
// across them; the column scalars are uniform over the work-group.
constexpr int kEpiVec = 8;

// sum_j s_jc I_jc of one row: the integer partials of the column's slices (columns c, k + c
// and the third-slice column) added over the splits, then scaled.  p points at the row in
// column 0 of split 0; ld = rows per column, sstride = one split.
__device__ __forceinline__ double i8_recombine(const int* __restrict__ p, size_t sstride,
                                               size_t splits, size_t ld, size_t k, size_t c,
                                               const I8Col& cc) {
  const int* a = p + c * ld;
  const int* b = p + (k + c) * ld;
  int i0 = 0, i1 = 0, i2 = 0;
  if (cc.c3 >= 0) {
    const int* d = p + (size_t)cc.c3 * ld;
    for (size_t sp = 0; sp < splits; ++sp) {
      i0 += a[sp * sstride];
      i1 += b[sp * sstride];
      i2 += d[sp * sstride];
    }
  } else {
    for (size_t sp = 0; sp < splits; ++sp) {
      i0 += a[sp * sstride];
      i1 += b[sp * sstride];
    }
  }
  double u = cc.v[1] * (double)i0 + cc.v[2] * (double)i1;
  if (cc.c3 >= 0) u += cc.v[3] * (double)i2;
  return u;
}

// Stage-1 epilogue: U'_mc = sum_j s_jc I_jmc over the column's slices;
// T_mc = s_m (U'_mc - 2p_m cs_c) + zs_m mean_c; W_mc = s_m T_mc (column-major, ldw).
// Grid (ceil(m / 256), ceil(k / kEpiVec)); a thread per marker row.
__global__ __launch_bounds__(256) void i8_stage1_epilogue_kernel(
    const int* __restrict__ part, size_t splits, size_t m_pad, size_t m, size_t k, size_t kp,
    const I8Col* __restrict__ cols, const double* __restrict__ sc,
    const double* __restrict__ p2, const double* __restrict__ zs, double* __restrict__ w,
    size_t ldw) {
  const size_t mm = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (mm >= m) return;
  const double s = sc[mm], pp = p2[mm], z = zs[mm];
  const size_t c0 = (size_t)blockIdx.y * kEpiVec, c1 = min(c0 + kEpiVec, k);
  for (size_t c = c0; c < c1; ++c) {
    const I8Col cc = cols[c];
    const double u = i8_recombine(part + mm, kp * m_pad, splits, m_pad, k, c, cc);
    const double t = s * (u - pp * cc.v[4]) + z * cc.v[0];
    w[mm + c * ldw] = s * t;
  }
}

// Stage-2 epilogue: M Y_ic = sum_j s_jc I_jic + nu_c (r_i - P2) - S2p_c; optional Sigma form.
// Grid (ceil(n / 256), ceil(k / kEpiVec)); a thread per sample row.
__global__ __launch_bounds__(256) void i8_stage2_epilogue_kernel(
    const int* __restrict__ part, size_t splits, size_t n_pad, size_t n, size_t k, size_t kp,
    const I8Col* __restrict__ cols, double inv_m, const double* __restrict__ rsum, double P2,
    double* __restrict__ out, size_t ldo, const double* __restrict__ v, size_t ldv,
    const double* __restrict__ wts, double tau0, double tau1) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double ri = rsum[i];
  const double wi = wts ? fmax(wts[i], 1e-30) : 1.0;
  const size_t c0 = (size_t)blockIdx.y * kEpiVec, c1 = min(c0 + kEpiVec, k);
  for (size_t c = c0; c < c1; ++c) {
    const I8Col cc = cols[c];
    double y = i8_recombine(part + i, kp * n_pad, splits, n_pad, k, c, cc);
    y = (y + cc.v[0] * (ri - P2) - cc.v[4]) * inv_m;
    if (wts) y = tau0 * v[i + c * ldv] / wi + tau1 * y;
    out[i + c * ldo] = y;
  }
}

// Work-groups of i8_gemm_kernel<KT, NS> resident at once on the device (cached per
// instantiation).
template <int KT, int NS>
static size_t resident_slots() {
  static size_t slots = 0;
  if (!slots) {
    int dev = 0, cus = 0, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, i8_gemm_kernel<KT, NS>, 512, 0);
    slots = (size_t)std::max(1, cus) * (size_t)std::max(1, occ);
  }
  return slots;
}

// Reduction splits: the fewest rounds of resident work-groups times chunks per split, with
// a 2% charge per extra split for the partial-sum traffic.  blocks = work-groups per split.
static size_t pick_splits(size_t blocks, size_t nchunks, size_t slots) {
  size_t best = 1;
  double best_cost = 1e300;
  for (size_t sp = 1; sp <= std::min<size_t>(nchunks, 32); ++sp) {
    const size_t per = ceil_div(nchunks, sp);
    const size_t eff = ceil_div(nchunks, per);
    const double rounds = (double)ceil_div(blocks * eff, slots);
    const double cost = rounds * (double)per * (1.0 + 0.02 * (double)(eff - 1));
    if (cost < best_cost - 1e-9) { best_cost = cost; best = eff; }
  }
  return best;
}

template <int KT>
static int i8_gemm(saige_grm* h, const uint8_t* codes, const int8_t* bq, size_t b_stride,
                   size_t rows_pad, size_t red_len, size_t groups, DevBuf& part,
                   size_t* splits_out, size_t vectors) {
  constexpr int NS = I8Ring<KT>::NS;
  SAIGE_ARG(rows_pad % 512 == 0 && red_len % kI8Chunk == 0 && b_stride >= red_len,
            "int8 GEMM shape does not match the blocked genotype layout");
  hipStream_t s = h->stream;
  const size_t kpt = (size_t)16 * KT * groups;
  const size_t rblk = rows_pad / 512;
  const size_t nkc = red_len / kI8Chunk;
  const size_t splits = pick_splits(rblk * groups, nkc, resident_slots<KT, NS>());
  const size_t cps = ceil_div(nkc, splits);
  const size_t sp = ceil_div(nkc, cps);
  SAIGE_TRY(part.ensure(sp * rows_pad * kpt * 4));
  const size_t blocks = round_up(rblk * sp, 8) * groups;
  SAIGE_ARG(blocks < (1u << 31), "int8 GEMM grid too large");
  SAIGE_TRY(timed_launch(h->timer, LaunchTag{(uint32_t)kpt, (uint32_t)vectors},
                         i8_gemm_kernel<KT, NS>, dim3((unsigned)blocks), dim3(512), s, codes, nkc,
                         bq, b_stride, cps, part.as<int>(), rows_pad, kpt, rblk, sp, groups));
  *splits_out = sp;
  return SAIGE_OK;
}

static int i8_gemm_dispatch(int kt, saige_grm* h, const uint8_t* codes, const int8_t* bq,
                            size_t b_stride, size_t rows_pad, size_t red_len, size_t groups,
                            DevBuf& part, size_t* splits, size_t vectors) {
  switch (kt) {
    case 1: return i8_gemm<1>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 2: return i8_gemm<2>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 3: return i8_gemm<3>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 4: return i8_gemm<4>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 5: return i8_gemm<5>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 6: return i8_gemm<6>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 7: return i8_gemm<7>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 8: return i8_gemm<8>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    case 9: return i8_gemm<9>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
    default: return i8_gemm<10>(h, codes, bq, b_stride, rows_pad, red_len, groups, part, splits, vectors);
  }
}

// Column groups for `tiles` 16-column tiles: one group of all of them up to kI8MaxKT; beyond,
// the group width (5 to kI8MaxKT tiles) with the fewest executed tiles, then the fewest
// groups.
constexpr int kI8MaxKT = 10;
static void i8_groups(size_t tiles, int* kt, size_t* groups) {
  if (tiles <= (size_t)kI8MaxKT) {
    *kt = (int)tiles;
    *groups = 1;
    return;
  }
  size_t best_cost = SIZE_MAX, best_g = 0;
  int best_kt = kI8MaxKT;
  for (int t = kI8MaxKT; t >= 5; --t) {
    const size_t g = ceil_div(tiles, (size_t)t);
    const size_t cost = g * (size_t)t;
    if (cost < best_cost || (cost == best_cost && g < best_g)) {
      best_cost = cost;
      best_g = g;
      best_kt = t;
    }
  }
  *kt = best_kt;
  *groups = best_g;
}

// One launch set: k vectors, those in sm with a third slice (k3 of them).
static int int8_product_launch(saige_grm* h, const SliceMap& sm, size_t k3, const double* d_v,
                               size_t ldv, size_t k, double* d_out, size_t ldo,
                               const double* d_w, const double* tau, size_t c0) {
  GrmDevice& g = h->g;
  MixedState* ms = mstate(h);
  hipStream_t s = h->stream;
  const size_t n = g.n, m = g.m, n_pad = g.n_pad, m_pad = g.m_pad;
  const size_t nch_n = std::max<size_t>(1, ceil_div(n, kDotChunk));
  const size_t nch_m = std::max<size_t>(1, ceil_div(m, kDotChunk));
  const size_t cols = 2 * k + k3;
  int kt = 0;
  size_t groups = 0;
  i8_groups(ceil_div(cols, 16), &kt, &groups);
  const size_t kpt = (size_t)16 * kt * groups;
  const unsigned pad_cols = (unsigned)(kpt - cols);
  SAIGE_TRY(ms->vq.ensure(kpt * n_pad));
  SAIGE_TRY(ms->wq.ensure(kpt * m_pad));
  SAIGE_TRY(ms->vs.ensure(n * k * 8));
  SAIGE_TRY(ms->ws_.ensure(m * k * 8));
  SAIGE_TRY(ms->w.ensure(m * k * 8));
  SAIGE_TRY(ms->red.ensure(4 * std::max(nch_n, nch_m) * k * 8));
  SAIGE_TRY(ms->cols.ensure(k * sizeof(I8Col) + k * 4 * 8));
  I8Col* colsc = ms->cols.as<I8Col>();
  double* ms4 = reinterpret_cast<double*>(ms->cols.as<char>() + k * sizeof(I8Col));
  const unsigned cb = (unsigned)ceil_div(k, 64);

  // Eight launches per product.  Column scalars (means, slice scales, sums of the sliced
  // values) stay as fixed-order partials per kDotChunk rows and are finished by the kernel
  // that needs them.
  double* psum = ms->red.as<double>();                // column sums of V, then of W
  double* pmax = psum + std::max(nch_n, nch_m) * k;   // partial maxima
  double* pmin = pmax + std::max(nch_n, nch_m) * k;   // partial minima
  double* pdot = pmin + std::max(nch_n, nch_m) * k;   // sums of the sliced values
  const dim3 gq_n((unsigned)ceil_div(n_pad, kDotChunk), (unsigned)k + pad_cols);
  const dim3 gq_m((unsigned)ceil_div(m_pad, kDotChunk), (unsigned)k + pad_cols);
  const dim3 ge1((unsigned)ceil_div(m, 256), (unsigned)ceil_div(k, kEpiVec));
  const dim3 ge2((unsigned)ceil_div(n, 256), (unsigned)ceil_div(k, kEpiVec));

  // ---- V: column sums, maxima and minima, slices + sums of the sliced values ----
  col_stats_kernel<<<dim3((unsigned)nch_n, (unsigned)k), kDotBlock, 0, s>>>(d_v, ldv, n, nch_n,
                                                                             psum, pmax, pmin);
  col_mean_scales_kernel<<<cb, 64, 0, s>>>(k, psum, pmax, pmin, 1.0 / (double)n, nch_n, ms4);
  quantize_slices_kernel<<<gq_n, kDotBlock, 0, s>>>(d_v, ldv, n, k, sm, k3, ms4, nch_n, n_pad,
                                                    ms->vq.as<int8_t>(), ms->vs.as<double>(), n,
                                                    nullptr, pdot);

  // ---- stage 1: I = G^T Q_v (rows = markers, reduction over samples) ----
  size_t splits1 = 0;
  SAIGE_TRY(i8_gemm_dispatch(kt, h, g.geno, ms->vq.as<int8_t>(), n_pad, m_pad, n_pad, groups,
                             ms->ipart1, &splits1, k));
  i8_col_scalars_kernel<<<cb, 64, 0, s>>>(k, sm, ms4, pdot, nch_n, colsc);
  i8_stage1_epilogue_kernel<<<ge1, 256, 0, s>>>(
      ms->ipart1.as<int>(), splits1, m_pad, m, k, kpt, colsc, ms->sc.as<double>(),
      ms->p2.as<double>(), ms->zs.as<double>(), ms->w.as<double>(), m);
  SAIGE_TRY(missing_stage1(h, ms->vs.as<double>(), n, k, nullptr, psum, nch_n, ms->w.as<double>(), m));
  LocoCodes lc;
  if (loco_codes(h, c0, k, lc)) loco_mask(h, ms->w.as<double>(), 1, m, m, k, lc, s);

  // ---- W: column sums over real markers, maxima and minima, slices + S2p partials ----
  col_stats_kernel<<<dim3((unsigned)nch_m, (unsigned)k), kDotBlock, 0, s>>>(
      ms->w.as<double>(), m, m, nch_m, psum, pmax, pmin);
  col_mean_scales_kernel<<<cb, 64, 0, s>>>(k, psum, pmax, pmin, 1.0 / (double)m, nch_m, ms4);
  quantize_slices_kernel<<<gq_m, kDotBlock, 0, s>>>(ms->w.as<double>(), m, m, k, sm, k3, ms4,
                                                    nch_m, m_pad, ms->wq.as<int8_t>(),
                                                    ms->ws_.as<double>(), m, ms->p2.as<double>(),
                                                    pdot);

  // ---- stage 2: I = G Q_w (rows = samples, reduction over markers) ----
  size_t splits2 = 0;
  SAIGE_TRY(i8_gemm_dispatch(kt, h, ms->geno_t.as<uint8_t>(), ms->wq.as<int8_t>(), m_pad, n_pad,
                             m_pad, groups, ms->ipart2, &splits2, k));
  i8_col_scalars_kernel<<<cb, 64, 0, s>>>(k, sm, ms4, pdot, nch_m, colsc);
  i8_stage2_epilogue_kernel<<<ge2, 256, 0, s>>>(
      ms->ipart2.as<int>(), splits2, n_pad, n, k, kpt, colsc, 1.0 / (double)m,
      ms->rsum.as<double>(), ms->P2, d_out, ldo, d_v, ldv, d_w, tau ? tau[0] : 0.0,
      tau ? tau[1] : 0.0);
  SAIGE_HIP(hipGetLastError());
  SAIGE_TRY(missing_stage2(h, ms->ws_.as<double>(), m, k, nullptr, psum, nch_m, d_out, ldo,
                           (tau ? tau[1] : 1.0) / (double)m));
  h->products += k;
  return SAIGE_OK;
}

int grm_int8_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out,
                     size_t ldo, const double* d_w, const double* tau) {
  SAIGE_ARG(h->g.n_pad <= kI8MaxRed && h->g.m_pad <= kI8MaxRed,
            "SAIGE_PREC_INT8 and SAIGE_PREC_INT8X2 need n and M <= 8,000,000 (exact int32 "
            "accumulation)");
  const int base = h->precision == SAIGE_PREC_INT8X2 ? 2 : 3;
  auto slices = [&](size_t c) {
    const int want = h->slice_req ? (int)h->slice_req[c] : 0;
    return std::max(base, want >= 3 ? 3 : 2);
  };
  // launches of at most kI8MaxVec vectors and kI8MaxCols slice columns
  for (size_t c0 = 0; c0 < k;) {
    SliceMap sm{};
    size_t kc = 0, k3 = 0;
    while (c0 + kc < k && kc < (size_t)kI8MaxVec) {
      const int sl = slices(c0 + kc);
      if (2 * (kc + 1) + k3 + (sl == 3) > (size_t)kI8MaxCols) break;
      if (sl == 3) {
        sm.bits[kc >> 5] |= 1u << (kc & 31);
        ++k3;
      }
      ++kc;
    }
    SAIGE_TRY(int8_product_launch(h, sm, k3, d_v + c0 * ldv, ldv, kc, d_out + c0 * ldo, ldo, d_w,
                                  tau, c0));
    c0 += kc;
  }
  return SAIGE_OK;
}

}  // namespace saige
