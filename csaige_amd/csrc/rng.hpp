// Random draws of the saige-rs fit-null path, reproduced from the published algorithms of
// the crates it uses (rand 0.8.5, rand_chacha 0.3.1, rand_core 0.6.4; saige-rs/Cargo.lock):
//   ChaCha8Rng::seed_from_u64(seed)     ai_reml.rs:123, fit_null.rs:234, variance_ratio.rs:113
//   rng.gen::<bool>()                   ai_reml.rs:127   (sign bit of next_u32)
//   rng.gen_range(0..n) for usize       fit_null.rs:237  (64-bit widening-multiply rejection)
//   indices.shuffle(&mut rng)           variance_ratio.rs:115 (Fisher-Yates from the end)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace saige {

// One 64-byte ChaCha block with 8 rounds, 64-bit block counter and stream 0 (the
// rand_chacha 0.3.1 layout); the same code serves the host stream and the device twin.
__host__ __device__ inline void chacha8_block(const uint32_t key[8], uint64_t ctr, uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2],
                    key[3], key[4], key[5], key[6], key[7], (uint32_t)ctr, (uint32_t)(ctr >> 32),
                    0u, 0u};
  uint32_t x[16];
  for (int i = 0; i < 16; ++i) x[i] = s[i];
#define SAIGE_RL(v, r) (((v) << (r)) | ((v) >> (32 - (r))))
#define SAIGE_QR(a, b, c, d)                                  \
  x[a] += x[b]; x[d] = SAIGE_RL(x[d] ^ x[a], 16);             \
  x[c] += x[d]; x[b] = SAIGE_RL(x[b] ^ x[c], 12);             \
  x[a] += x[b]; x[d] = SAIGE_RL(x[d] ^ x[a], 8);              \
  x[c] += x[d]; x[b] = SAIGE_RL(x[b] ^ x[c], 7);
  for (int dr = 0; dr < 4; ++dr) {
    SAIGE_QR(0, 4, 8, 12) SAIGE_QR(1, 5, 9, 13) SAIGE_QR(2, 6, 10, 14) SAIGE_QR(3, 7, 11, 15)
    SAIGE_QR(0, 5, 10, 15) SAIGE_QR(1, 6, 11, 12) SAIGE_QR(2, 7, 8, 13) SAIGE_QR(3, 4, 9, 14)
  }
#undef SAIGE_QR
#undef SAIGE_RL
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

class ChaCha8 {
 public:
  explicit ChaCha8(uint64_t seed) {
    // rand_core 0.6.4 seed_from_u64: a PCG32 stream fills the 32-byte key.
    uint64_t st = seed;
    for (int w = 0; w < 8; ++w) {
      st = st * 6364136223846793005ull + 11634580027462260723ull;
      uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
      uint32_t rot = (uint32_t)(st >> 59);
      key_[w] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
  }
  uint32_t next_u32() {
    if (pos_ == 64) refill();
    return buf_[pos_++];
  }
  uint64_t next_u64() {  // two consecutive stream words, low first
    uint64_t lo = next_u32();
    return lo | ((uint64_t)next_u32() << 32);
  }
  bool gen_bool() { return (next_u32() >> 31) != 0; }
  uint64_t gen_range64(uint64_t n) {
    const uint64_t zone = (n << __builtin_clzll(n)) - 1;
    for (;;) {
      unsigned __int128 m = (unsigned __int128)next_u64() * n;
      if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
    }
  }
  uint32_t gen_range32(uint32_t n) {
    const uint32_t zone = (n << __builtin_clz(n)) - 1;
    for (;;) {
      uint64_t m = (uint64_t)next_u32() * n;
      if ((uint32_t)m <= zone) return (uint32_t)(m >> 32);
    }
  }
  template <class T>
  void shuffle(std::vector<T>& v) {
    for (size_t i = v.size(); i > 1; --i) {
      size_t j = (i <= 0xffffffffull) ? gen_range32((uint32_t)i) : (size_t)gen_range64(i);
      std::swap(v[i - 1], v[j]);
    }
  }

 private:
  // Four consecutive 64-byte ChaCha blocks.
  void refill() {
    for (int b = 0; b < 4; ++b) chacha8_block(key_, ctr_ + b, buf_ + 16 * b);
    ctr_ += 4;
    pos_ = 0;
  }

 public:
  const uint32_t* key() const { return key_; }

 private:
  uint32_t key_[8];
  uint32_t buf_[64];
  int pos_ = 64;
  uint64_t ctr_ = 0;
};

}  // namespace saige
