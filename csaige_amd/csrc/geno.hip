// Genotype decode, allele counts and the resident packed-genotype build.
//
// Reference: saige-geno/src/plink.rs:170-267 (decode_genotype, read_marker_raw,
// set_sample_subset, read_marker) and traits.rs:42-54 (MarkerData::compute_af).
#include "common.hpp"

#include <atomic>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

namespace saige {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
void clear_error() { g_err.clear(); }

// ---- device memory cache (see common.hpp) ----
namespace {
struct MemCache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void*>> blocks;  // device -> size -> free block
  size_t cached = 0;
};
MemCache& mem_cache() {
  static MemCache* c = new MemCache();  // never destroyed: blocks may be returned at exit
  return *c;
}
constexpr size_t kCacheLimit = size_t(32) << 30;  // cached bytes kept at most
size_t alloc_round(size_t b) {
  return b < (size_t(64) << 20) ? round_up(b, size_t(64) << 10) : round_up(b, size_t(2) << 20);
}
// frees the cached blocks of `device` (all devices if < 0); caller holds the lock
void cache_release_locked(MemCache& c, int device) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& kv : c.blocks) {
    if (device >= 0 && kv.first != device) continue;
    (void)hipSetDevice(kv.first);
    for (auto& b : kv.second) {
      (void)hipFree(b.second);
      c.cached -= b.first;
    }
    kv.second.clear();
  }
  (void)hipSetDevice(cur);
}
}  // namespace

static std::atomic<uint64_t> g_devbuf_generation{0};
uint64_t devbuf_generation() { return g_devbuf_generation.load(std::memory_order_acquire); }
void devbuf_generation_bump() { g_devbuf_generation.fetch_add(1, std::memory_order_acq_rel); }

void* cache_alloc(size_t bytes, size_t* rounded, int* device) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const size_t r = alloc_round(bytes);
  MemCache& c = mem_cache();
  void* reuse = nullptr;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto& m = c.blocks[dev];
    auto it = m.lower_bound(r);  // smallest free block that fits, if not more than twice as big
    if (it != m.end() && it->first <= 2 * r) {
      reuse = it->second;
      *rounded = it->first;
      c.cached -= it->first;
      m.erase(it);
    }
  }
  if (reuse) {
    // work enqueued on any stream before the block was freed may still read or write it
    (void)hipDeviceSynchronize();
    *device = dev;
    return reuse;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(c.mu);
      cache_release_locked(c, dev);
    }
    e = hipMalloc(&p, r);
  }
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", r, hipGetErrorString(e));
    return nullptr;
  }
  *rounded = r;
  *device = dev;
  return p;
}

void cache_free(void* p, size_t rounded, int device) {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.blocks[device].emplace(rounded, p);
  c.cached += rounded;
  while (c.cached > kCacheLimit) {  // drop the largest cached blocks first
    int bd = -1;
    size_t bs = 0;
    for (auto& kv : c.blocks)
      if (!kv.second.empty() && kv.second.rbegin()->first >= bs) {
        bs = kv.second.rbegin()->first;
        bd = kv.first;
      }
    if (bd < 0) break;
    auto it = std::prev(c.blocks[bd].end());
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(bd);
    (void)hipFree(it->second);
    (void)hipSetDevice(cur);
    c.cached -= it->first;
    c.blocks[bd].erase(it);
  }
}

int use_device(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return SAIGE_ERR_DEVICE;
  }
  if (device < 0 || device >= count) {
    set_error("device %d out of range (%d devices)", device, count);
    return SAIGE_ERR_ARG;
  }
  SAIGE_HIP(hipSetDevice(device));
  return SAIGE_OK;
}

// One 2-bit code of sample s (model order -> file order through idx).
__device__ __forceinline__ uint32_t code_at(const uint8_t* row, const uint32_t* idx, size_t k) {
  size_t s = idx ? idx[k] : k;
  return (row[s >> 2] >> (2 * (s & 3))) & 3u;
}

// Per marker: allele-1 count (sum of dosages over non-missing samples) and n_valid.
// The sum of the integer dosages is exact, so ac / af / mac match compute_af bit for bit.
__global__ void marker_stats_kernel(const uint8_t* __restrict__ bed, size_t bpm,
                                    const uint32_t* __restrict__ idx, size_t n, size_t m,
                                    double* __restrict__ ac, uint32_t* __restrict__ nvalid,
                                    double* __restrict__ af, double* __restrict__ mac) {
  size_t j = blockIdx.x;
  if (j >= m) return;
  const uint8_t* row = bed + j * bpm;
  uint32_t cnt = 0, sum = 0;
  for (size_t k = threadIdx.x; k < n; k += blockDim.x) {
    uint32_t c = code_at(row, idx, k);
    // 00 -> 2, 10 -> 1, 11 -> 0, 01 -> missing
    cnt += (c != 1u);
    sum += (c == 0u) ? 2u : (c == 2u ? 1u : 0u);
  }
  __shared__ uint32_t s_cnt[256], s_sum[256];
  s_cnt[threadIdx.x] = cnt;
  s_sum[threadIdx.x] = sum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + w];
      s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double s = (double)s_sum[0];
    double k = (double)s_cnt[0];
    ac[j] = s;
    nvalid[j] = s_cnt[0];
    af[j] = s_cnt[0] > 0 ? s / (2.0 * k) : 0.0;
    mac[j] = fmin(s, 2.0 * k - s);
  }
}

// Same counts without a sample subset: 16 codes per 32-bit word.  With lo / hi the low and
// high bits of each code: missing = lo & ~hi (01), dosage 2 = ~lo & ~hi (00), 1 = hi & ~lo (10).
__global__ void marker_stats_packed_kernel(const uint8_t* __restrict__ bed, size_t bpm, size_t n,
                                           size_t m, double* __restrict__ ac,
                                           uint32_t* __restrict__ nvalid, double* __restrict__ af,
                                           double* __restrict__ mac) {
  size_t j = blockIdx.x;
  if (j >= m) return;
  const uint8_t* row = bed + j * bpm;
  const size_t nwords = n / 16;
  uint32_t miss = 0, sum = 0;
  const bool aligned = ((reinterpret_cast<uintptr_t>(row) & 3u) == 0);
  for (size_t k = threadIdx.x; k < nwords; k += blockDim.x) {
    uint32_t w;
    if (aligned) w = reinterpret_cast<const uint32_t*>(row)[k];
    else w = (uint32_t)row[4 * k] | ((uint32_t)row[4 * k + 1] << 8) | ((uint32_t)row[4 * k + 2] << 16) |
             ((uint32_t)row[4 * k + 3] << 24);
    const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
    miss += __popc(lo & ~hi);
    sum += 2u * __popc(~lo & ~hi & 0x55555555u) + __popc(hi & ~lo);
  }
  for (size_t i = nwords * 16 + threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t c = (row[i >> 2] >> (2 * (i & 3))) & 3u;
    miss += (c == 1u);
    sum += (c == 0u) ? 2u : (c == 2u ? 1u : 0u);
  }
  __shared__ uint32_t s_m[256], s_s[256];
  s_m[threadIdx.x] = miss;
  s_s[threadIdx.x] = sum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_m[threadIdx.x] += s_m[threadIdx.x + w];
      s_s[threadIdx.x] += s_s[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t cnt = (uint32_t)n - s_m[0];
    const double sd = (double)s_s[0], k = (double)cnt;
    ac[j] = sd;
    nvalid[j] = cnt;
    af[j] = cnt > 0 ? sd / (2.0 * k) : 0.0;
    mac[j] = fmin(sd, 2.0 * k - sd);
  }
}

__global__ void decode_kernel(const uint8_t* __restrict__ bed, size_t bpm,
                              const uint32_t* __restrict__ idx, size_t n, size_t m,
                              double* __restrict__ out, size_t ldo) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t j = blockIdx.y;
  if (k >= n || j >= m) return;
  uint32_t c = code_at(bed + j * bpm, idx, k);
  out[j * ldo + k] = c == 0u ? 2.0 : (c == 2u ? 1.0 : (c == 3u ? 0.0 : __longlong_as_double(0x7ff8000000000000ll)));
}

// Repacking selected markers into padded SNP-major rows of the model-ordered samples:
// output byte b of a row holds model samples 4b..4b+3; samples >= n get code 01.  Source
// row j is midx[j] (or j) of a buffer with src_bpm bytes per row; sample k of the model
// order is source sample idx[k] (or k).  Every read stays inside [floor4(row start),
// row start + src_bpm).

// 16 samples of an output dword at or past sample n get code 01.
__device__ __forceinline__ uint32_t pad_codes(uint32_t w, long long valid) {
  if (valid >= 16) return w;
  const uint32_t keep = valid <= 0 ? 0u : ((1u << (2 * valid)) - 1u);
  return (w & keep) | (0x55555555u & ~keep);
}

// Identity sample order: output dword o of a row is source bytes 4o..4o+3, a funnel shift
// of two aligned source dwords; each thread writes 16 bytes.
__global__ __launch_bounds__(256) void repack_copy_kernel(
    const uint8_t* __restrict__ src, size_t src_bpm, const uint32_t* __restrict__ midx,
    size_t n, size_t m, uint8_t* __restrict__ dst, size_t row_bytes) {
  const size_t j = blockIdx.y;
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m || 16 * p >= row_bytes) return;
  const uint8_t* row = src + (size_t)(midx ? midx[j] : j) * src_bpm;
  const size_t off = (size_t)row & 3;
  const uint32_t* base = reinterpret_cast<const uint32_t*>(row - off);
  const size_t nbytes = (n + 3) / 4;
  uint32_t out[4];
  if (16 * p + 20 <= src_bpm) {  // base[4p .. 4p + 4] lie inside the row
    uint32_t d[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) d[e] = base[4 * p + e];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      out[e] = off ? __builtin_amdgcn_alignbyte(d[e + 1], d[e], (uint32_t)off) : d[e];
  } else {  // the row's last bytes, one at a time
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t w = 0;
      for (int b = 0; b < 4; ++b) {
        const size_t q = 16 * p + 4 * e + b;
        w |= (uint32_t)(q < nbytes ? row[q] : 0x55u) << (8 * b);
      }
      out[e] = w;
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) out[e] = pad_codes(out[e], (long long)n - (long long)(64 * p + 16 * e));
  *reinterpret_cast<uint4*>(dst + j * row_bytes + 16 * p) = make_uint4(out[0], out[1], out[2], out[3]);
}

constexpr int kRepackRows = 8;  // source rows staged per work-group at most

// A sample list: rows_per_group source rows staged in LDS (`stage` bytes apart, with
// aligned dword loads), then every output dword gathers its 16 codes from them, reading
// each list entry once for all the rows.
__global__ __launch_bounds__(256) void repack_stage_kernel(
    const uint8_t* __restrict__ src, size_t src_bpm, const uint32_t* __restrict__ midx,
    const uint32_t* __restrict__ idx, size_t n, size_t m, uint8_t* __restrict__ dst,
    size_t row_bytes, size_t stage, int rows_per_group) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  const size_t j0 = (size_t)blockIdx.x * rows_per_group;
  const int rows = (int)min((size_t)rows_per_group, m - j0);
  uint32_t offs[kRepackRows];
#pragma unroll
  for (int r = 0; r < kRepackRows; ++r) {
    offs[r] = 0;
    if (r < rows) {
      const uint8_t* row = src + (size_t)(midx ? midx[j0 + r] : j0 + r) * src_bpm;
      const size_t off = (size_t)row & 3;
      offs[r] = (uint32_t)off;
      const uint32_t* base = reinterpret_cast<const uint32_t*>(row - off);
      const size_t nd = (off + src_bpm + 3) / 4;  // dwords covering the row
      const size_t nfull = (off + src_bpm) / 4;   // those that end inside it
      uint32_t* l = lds + r * (stage / 4);
      for (size_t d = threadIdx.x; d < nd; d += 256) {
        uint32_t w = 0;
        if (d < nfull) {
          w = base[d];
        } else {
          for (int b = 0; b < 4; ++b) {
            const size_t q = 4 * d + b;
            if (q >= off && q - off < src_bpm) w |= (uint32_t)row[q - off] << (8 * b);
          }
        }
        l[d] = w;
      }
    }
  }
  __syncthreads();
  const uint8_t* lb = reinterpret_cast<const uint8_t*>(lds);
  for (size_t o = threadIdx.x; o < row_bytes / 4; o += 256) {
    uint32_t w[kRepackRows];
#pragma unroll
    for (int r = 0; r < kRepackRows; ++r) w[r] = 0;
    uint32_t sl[16];  // the 16 list entries of this output dword
    if (16 * o + 16 <= n) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint4 q = reinterpret_cast<const uint4*>(idx + 16 * o)[v];
        sl[4 * v] = q.x; sl[4 * v + 1] = q.y; sl[4 * v + 2] = q.z; sl[4 * v + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int f = 0; f < 16; ++f) sl[f] = (16 * o + f < n) ? idx[16 * o + f] : 0u;
    }
#pragma unroll
    for (int f = 0; f < 16; ++f) {
      const size_t k = 16 * o + f;
      if (k < n) {
        const uint32_t sidx = sl[f];
#pragma unroll
        for (int r = 0; r < kRepackRows; ++r)
          if (r < rows)
            w[r] |= ((uint32_t)(lb[r * stage + offs[r] + (sidx >> 2)] >> (2 * (sidx & 3))) & 3u)
                    << (2 * f);
      } else {
#pragma unroll
        for (int r = 0; r < kRepackRows; ++r) w[r] |= 1u << (2 * f);
      }
    }
#pragma unroll
    for (int r = 0; r < kRepackRows; ++r)
      if (r < rows) *reinterpret_cast<uint32_t*>(dst + (j0 + r) * row_bytes + 4 * o) = w[r];
  }
}

// A sample list with rows too long to stage: each output dword gathers its 16 codes from
// global memory.
__global__ __launch_bounds__(256) void repack_gather_kernel(
    const uint8_t* __restrict__ src, size_t src_bpm, const uint32_t* __restrict__ midx,
    const uint32_t* __restrict__ idx, size_t n, size_t m, uint8_t* __restrict__ dst,
    size_t row_bytes) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t j = blockIdx.y;
  if (j >= m || 4 * o >= row_bytes) return;
  const uint8_t* row = src + (size_t)(midx ? midx[j] : j) * src_bpm;
  uint32_t w = 0;
  for (int f = 0; f < 16; ++f) {
    const size_t k = 16 * o + f;
    uint32_t c = 1u;
    if (k < n) {
      const uint32_t sidx = idx[k];
      c = (row[sidx >> 2] >> (2 * (sidx & 3))) & 3u;
    }
    w |= c << (2 * f);
  }
  *reinterpret_cast<uint32_t*>(dst + j * row_bytes + 4 * o) = w;
}

// Launches the repack of m rows (row_bytes a multiple of 16).
static int repack_rows(hipStream_t s, const uint8_t* src, size_t bpm, const uint32_t* d_midx,
                       const uint32_t* d_idx, size_t n, size_t m, uint8_t* dst,
                       size_t row_bytes) {
  const size_t ychunk = 65535;
  if (!d_idx) {
    for (size_t j0 = 0; j0 < m; j0 += ychunk) {
      const size_t mc = std::min(ychunk, m - j0);
      repack_copy_kernel<<<dim3((unsigned)ceil_div(row_bytes / 16, 256), (unsigned)mc), 256, 0, s>>>(
          d_midx ? src : src + j0 * bpm, bpm, d_midx ? d_midx + j0 : nullptr, n, mc,
          dst + j0 * row_bytes, row_bytes);
    }
    SAIGE_HIP(hipGetLastError());
    return SAIGE_OK;
  }
  const size_t stage = round_up(bpm + 8, 16);  // covers floor4 offset + row, in dwords
  const size_t rows = std::min<size_t>(kRepackRows, 65536 / stage);
  if (rows >= 1) {
    repack_stage_kernel<<<(unsigned)ceil_div(m, rows), 256, rows * stage, s>>>(
        src, bpm, d_midx, d_idx, n, m, dst, row_bytes, stage, (int)rows);
  } else {
    for (size_t j0 = 0; j0 < m; j0 += ychunk) {
      const size_t mc = std::min(ychunk, m - j0);
      repack_gather_kernel<<<dim3((unsigned)ceil_div(row_bytes / 4, 256), (unsigned)mc), 256, 0, s>>>(
          d_midx ? src : src + j0 * bpm, bpm, d_midx ? d_midx + j0 : nullptr, d_idx, n, mc,
          dst + j0 * row_bytes, row_bytes);
    }
  }
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

int repack_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                       size_t n, const uint32_t* marker_idx, size_t m, uint8_t* d_dst,
                       size_t row_bytes) {
  if (m == 0) return SAIGE_OK;
  DevBuf d_mi;
  if (marker_idx) {
    SAIGE_TRY(d_mi.ensure(m * 4));
    SAIGE_HIP(hipMemcpyAsync(d_mi.p, marker_idx, m * 4, hipMemcpyHostToDevice, s));
  }
  SAIGE_TRY(repack_rows(s, d_bed, bpm, marker_idx ? d_mi.as<uint32_t>() : nullptr, d_idx, n, m,
                        d_dst, row_bytes));
  SAIGE_HIP(hipStreamSynchronize(s));
  return SAIGE_OK;
}

// Allele counts of m markers of device-resident packed rows, into device memory.
int marker_ac_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, size_t n, size_t m,
                          double* d_ac) {
  if (m == 0) return SAIGE_OK;
  DevBuf d_nv, d_af, d_mac;
  SAIGE_TRY(d_nv.ensure(m * 4));
  SAIGE_TRY(d_af.ensure(m * 8));
  SAIGE_TRY(d_mac.ensure(m * 8));
  const size_t chunk = 1u << 30;
  for (size_t j0 = 0; j0 < m; j0 += chunk) {
    const size_t mc = std::min(chunk, m - j0);
    marker_stats_packed_kernel<<<(unsigned)mc, 256, 0, s>>>(
        d_bed + j0 * bpm, bpm, n, mc, d_ac + j0, d_nv.as<uint32_t>() + j0,
        d_af.as<double>() + j0, d_mac.as<double>() + j0);
  }
  SAIGE_HIP(hipGetLastError());
  SAIGE_HIP(hipStreamSynchronize(s));  // the scratch outputs are released on return
  return SAIGE_OK;
}

// Allele counts of m markers of a device-resident .bed body (host outputs).
int marker_stats_from_device(hipStream_t s, const uint8_t* d_bed, size_t bpm, const uint32_t* d_idx,
                             size_t n, size_t m, double* ac, uint32_t* n_valid, double* af,
                             double* mac) {
  if (m == 0) return SAIGE_OK;
  DevBuf d_ac, d_nv, d_af, d_mac;
  SAIGE_TRY(d_ac.ensure(m * 8));
  SAIGE_TRY(d_nv.ensure(m * 4));
  SAIGE_TRY(d_af.ensure(m * 8));
  SAIGE_TRY(d_mac.ensure(m * 8));
  const size_t chunk = 1u << 30;
  for (size_t j0 = 0; j0 < m; j0 += chunk) {
    size_t mc = std::min(chunk, m - j0);
    if (d_idx)
      marker_stats_kernel<<<(unsigned)mc, 256, 0, s>>>(d_bed + j0 * bpm, bpm, d_idx, n, mc,
                                                       d_ac.as<double>() + j0, d_nv.as<uint32_t>() + j0,
                                                       d_af.as<double>() + j0, d_mac.as<double>() + j0);
    else
      marker_stats_packed_kernel<<<(unsigned)mc, 256, 0, s>>>(
          d_bed + j0 * bpm, bpm, n, mc, d_ac.as<double>() + j0, d_nv.as<uint32_t>() + j0,
          d_af.as<double>() + j0, d_mac.as<double>() + j0);
    SAIGE_HIP(hipGetLastError());
  }
  SAIGE_HIP(hipMemcpyAsync(ac, d_ac.p, m * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipMemcpyAsync(n_valid, d_nv.p, m * 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipMemcpyAsync(af, d_af.p, m * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipMemcpyAsync(mac, d_mac.p, m * 8, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  return SAIGE_OK;
}

// Upload the rows of the selected markers (marker_idx[j], or j) in chunks and repack.
// dst: device rows of row_bytes.  Returns status.
int upload_repack(hipStream_t s, const uint8_t* bed_body, size_t bpm, size_t n_total,
                  const uint32_t* d_idx, size_t n, const uint32_t* marker_idx, size_t m,
                  uint8_t* d_dst, size_t row_bytes) {
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(m, (size_t)(256u << 20) / std::max<size_t>(bpm, 1)));
  DevBuf staging;
  SAIGE_TRY(staging.ensure(chunk * bpm));
  std::vector<uint8_t> host;
  bool contiguous = (marker_idx == nullptr);
  if (!contiguous) host.resize(chunk * bpm);
  (void)n_total;
  for (size_t j0 = 0; j0 < m; j0 += chunk) {
    size_t mc = std::min(chunk, m - j0);
    const uint8_t* srcp;
    if (contiguous) {
      srcp = bed_body + j0 * bpm;
    } else {
      for (size_t j = 0; j < mc; ++j)
        memcpy(host.data() + j * bpm, bed_body + (size_t)marker_idx[j0 + j] * bpm, bpm);
      srcp = host.data();
    }
    SAIGE_HIP(hipMemcpyAsync(staging.p, srcp, mc * bpm, hipMemcpyHostToDevice, s));
    SAIGE_TRY(repack_rows(s, staging.as<uint8_t>(), bpm, nullptr, d_idx, n, mc,
                          d_dst + j0 * row_bytes, row_bytes));
    SAIGE_HIP(hipStreamSynchronize(s));
  }
  return SAIGE_OK;
}

// The index list is uploaded unless it is the identity over all samples (then the kernels
// take their faster no-subset paths; the result is the same).
static int upload_idx(hipStream_t s, const uint32_t* sample_idx, size_t n, size_t n_total,
                      DevBuf& d) {
  if (!sample_idx) return SAIGE_OK;
  if (n == n_total) {
    bool ident = true;
    for (size_t k = 0; k < n && ident; ++k) ident = (sample_idx[k] == k);
    if (ident) return SAIGE_OK;
  }
  for (size_t k = 0; k < n; ++k)
    if (sample_idx[k] >= n_total) {
      set_error("sample_idx[%zu] = %u >= n_total %zu", k, sample_idx[k], n_total);
      return SAIGE_ERR_ARG;
    }
  SAIGE_TRY(d.ensure(n * sizeof(uint32_t)));
  SAIGE_HIP(hipMemcpyAsync(d.p, sample_idx, n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  return SAIGE_OK;
}

int upload_sample_idx(hipStream_t s, const uint32_t* sample_idx, size_t n, size_t n_total,
                      DevBuf& d) {
  return upload_idx(s, sample_idx, n, n_total, d);
}

}  // namespace saige

using namespace saige;

extern "C" {

const char* saige_last_error(void) { return g_err.c_str(); }

void saige_release_cached_memory(void) {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  cache_release_locked(c, -1);
}
const char* saige_version(void) { return "saige_gpu 0.1 (gfx950)"; }

int saige_plink_marker_stats(int device, const uint8_t* bed_body, size_t bpm, size_t n_total,
                             size_t m, const uint32_t* sample_idx, size_t n, double* ac,
                             uint32_t* n_valid, double* af, double* mac) {
  clear_error();
  SAIGE_ARG(bed_body && ac && n_valid && af && mac, "null pointer");
  SAIGE_ARG(bpm >= (n_total + 3) / 4, "bytes_per_marker < ceil(n_total/4)");
  if (!sample_idx) SAIGE_ARG(n == n_total, "n must equal n_total without a subset");
  if (m == 0) return SAIGE_OK;
  SAIGE_TRY(use_device(device));
  hipStream_t s;
  SAIGE_HIP(hipStreamCreate(&s));
  int rc = SAIGE_OK;
  {
    DevBuf d_idx, d_bed, d_ac, d_nv, d_af, d_mac;
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(m, (size_t)(512u << 20) / bpm));
    rc = upload_idx(s, sample_idx, n, n_total, d_idx);
    if (rc == SAIGE_OK) rc = d_bed.ensure(chunk * bpm);
    if (rc == SAIGE_OK) rc = d_ac.ensure(chunk * 8);
    if (rc == SAIGE_OK) rc = d_nv.ensure(chunk * 4);
    if (rc == SAIGE_OK) rc = d_af.ensure(chunk * 8);
    if (rc == SAIGE_OK) rc = d_mac.ensure(chunk * 8);
    for (size_t j0 = 0; rc == SAIGE_OK && j0 < m; j0 += chunk) {
      size_t mc = std::min(chunk, m - j0);
      hipError_t e = hipMemcpyAsync(d_bed.p, bed_body + j0 * bpm, mc * bpm, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) {
        if (d_idx.p)
          marker_stats_kernel<<<(unsigned)mc, 256, 0, s>>>(d_bed.as<uint8_t>(), bpm,
                                                           d_idx.as<uint32_t>(), n, mc,
                                                           d_ac.as<double>(), d_nv.as<uint32_t>(),
                                                           d_af.as<double>(), d_mac.as<double>());
        else
          marker_stats_packed_kernel<<<(unsigned)mc, 256, 0, s>>>(
              d_bed.as<uint8_t>(), bpm, n, mc, d_ac.as<double>(), d_nv.as<uint32_t>(),
              d_af.as<double>(), d_mac.as<double>());
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(ac + j0, d_ac.p, mc * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(n_valid + j0, d_nv.p, mc * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(af + j0, d_af.p, mc * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(mac + j0, d_mac.p, mc * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) { set_error("marker stats: %s", hipGetErrorString(e)); rc = SAIGE_ERR_DEVICE; }
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

int saige_plink_decode(int device, const uint8_t* bed_body, size_t bpm, size_t n_total, size_t m,
                       const uint32_t* sample_idx, size_t n, double* dosages) {
  clear_error();
  SAIGE_ARG(bed_body && dosages, "null pointer");
  SAIGE_ARG(bpm >= (n_total + 3) / 4, "bytes_per_marker < ceil(n_total/4)");
  if (!sample_idx) SAIGE_ARG(n == n_total, "n must equal n_total without a subset");
  if (m == 0 || n == 0) return SAIGE_OK;
  SAIGE_TRY(use_device(device));
  hipStream_t s;
  SAIGE_HIP(hipStreamCreate(&s));
  int rc = SAIGE_OK;
  {
    DevBuf d_idx, d_bed, d_out;
    rc = upload_idx(s, sample_idx, n, n_total, d_idx);
    if (rc == SAIGE_OK) rc = d_bed.ensure(m * bpm);
    if (rc == SAIGE_OK) rc = d_out.ensure(m * n * 8);
    if (rc == SAIGE_OK) {
      hipError_t e = hipMemcpyAsync(d_bed.p, bed_body, m * bpm, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) {
        dim3 grid((unsigned)ceil_div(n, 256), (unsigned)m);
        decode_kernel<<<grid, 256, 0, s>>>(d_bed.as<uint8_t>(), bpm, d_idx.as<uint32_t>(), n, m,
                                           d_out.as<double>(), n);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(dosages, d_out.p, m * n * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) { set_error("decode: %s", hipGetErrorString(e)); rc = SAIGE_ERR_DEVICE; }
    }
  }
  (void)hipStreamDestroy(s);
  return rc;
}

}  // extern "C"
