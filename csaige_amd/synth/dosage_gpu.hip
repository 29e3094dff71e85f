// Device twin of synth_dosage_u8 (cohort.cpp): the step-2 workload's uint8 dosages generated
// in HBM from the same counter-based hashes, byte for byte equal to the host generator, so
// that a benchmark can stream 1,000,000 variants x 400,000 samples (400 GB) through the score
// test batch by batch without holding or copying them.  Synthetic-data harness, not part of
// the saige-rs path.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t h3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed * 0x9e3779b97f4a7c15ull + stream) ^ (idx + 0x632be59bd9b4e019ull));
}
constexpr uint64_t kDosage = 11, kNoise = 13;  // the host generator's stream ids

// Variant blockIdx.y of the batch (global index first + blockIdx.y), 4 samples per thread.
__global__ __launch_bounds__(256) void dosage_u8_kernel(uint64_t seed, size_t n, size_t first,
                                                        const double* __restrict__ mafs,
                                                        uint8_t* __restrict__ q, size_t ld) {
#pragma clang fp contract(off)
  const size_t vv = blockIdx.y;
  const size_t v = first + vv;
  const size_t i0 = 4 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i0 >= ld) return;
  const double maf = mafs[vv];
  const bool noisy = (v % 4) == 3;
  const uint64_t key = mix64(seed * 0x9e3779b97f4a7c15ull + kDosage);
  const uint64_t nkey = mix64(seed * 0x9e3779b97f4a7c15ull + kNoise);
  uint32_t word = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const size_t i = i0 + u;
    uint32_t qq = 0;
    if (i < n) {
      const uint64_t h = mix64(key ^ (v * n + i + 0x632be59bd9b4e019ull));
      const double u1 = (double)(uint32_t)h * (1.0 / 4294967296.0);
      const double u2 = (double)(uint32_t)(h >> 32) * (1.0 / 4294967296.0);
      const int c = (u1 < maf) + (u2 < maf);
      int r = 127 * c;
      if (noisy) {
        const uint64_t g = mix64(nkey ^ (v * n + i + 0x632be59bd9b4e019ull));
        const uint32_t s4 = (uint32_t)(g & 0xFFFFu) + (uint32_t)((g >> 16) & 0xFFFFu) +
                            (uint32_t)((g >> 32) & 0xFFFFu) + (uint32_t)(g >> 48);
        const double t = (double)s4 * (1.0 / 65536.0) - 2.0;
        const double e = (t * 1.7320508075688772) * 12.0;
        const double y = 127.0 * (double)c + e;
        r = (int)llround(y);
        r = r < 0 ? 0 : (r > 254 ? 254 : r);
      }
      qq = (uint32_t)r;
    }
    word |= qq << (8 * u);
  }
  uint8_t* row = q + vv * ld;
  if ((ld & 3) == 0) {
    *reinterpret_cast<uint32_t*>(row + i0) = word;
  } else {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u < ld) row[i0 + u] = (uint8_t)(word >> (8 * u));
  }
}

}  // namespace

extern "C" {

// Dosages of variants [first, first + nvar) into the device rows d_q (nvar x ld bytes, zero
// past n), given their MAFs (d_mafs, device, from synth_dosage_mafs), on `stream`.
// Returns 0, or 1 on a bad argument or launch error.  Requires ld >= n and the row start
// q (and ld) 4-byte aligned for the word stores.
int synth_dosage_u8_device(uint64_t seed, size_t n, size_t first, size_t nvar, size_t ld,
                           const double* d_mafs, uint8_t* d_q, void* stream) {
  if (ld < n || nvar == 0 || nvar > 65535 || ((uintptr_t)d_q & 3)) return 1;
  const dim3 grid((unsigned)((ld + 1023) / 1024), (unsigned)nvar);
  dosage_u8_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(seed, n, first, d_mafs, d_q, ld);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
