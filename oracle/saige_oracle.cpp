// CPU ORACLE — test infrastructure only.
//
// A line-by-line C++ restatement of the saige-rs CPU path for the two hot paths
// (fit-null AI-REML with the on-the-fly GRM product, and the step-2 score test with
// variance ratio and saddle-point approximation).  Only the tests, the smoke check and
// the CPU-baseline leg of the benchmark load this library, and only as the checker or
// as the reported CPU baseline.  The product (libsaige_gpu.so) never links or calls it.
//
// Every function cites the saige-rs file:line it restates (paths relative to the
// reference checkout).  Floating-point order of operations follows the Rust source:
// sequential iterator sums, no fused multiply-add (compiled with -ffp-contract=off),
// left-associative products exactly as written.
//
// GRM precision switch: saige-rs stores standardized genotypes as f32 and accumulates
// the product in f32 (saige-core/src/glmm/pcg.rs:88-180).  ORC_GRM_F32 reproduces that
// (deterministic at one rayon thread: a single chunk, pcg.rs:141-145).  ORC_GRM_F64 is
// the same algorithm with f64 storage and accumulation: the "FP64 oracle" the plain GPU
// path is held to.
//
// Pinning (what checks each function; see tests/):
//   decode / compute_af  : reference KATs plink.rs:289-303 and the .frq / AC fixtures
//   GRM product          : dense brute force Z Z^T v / M (numpy) + symmetry/PSD invariants
//   PCG                  : KATs decomposition.rs:470-526, pcg.rs:202-223; numpy.linalg.solve
//   ChaCha8 core         : OpenSSL chacha20 keystream (20 rounds) through the same core
//   PCG32 (seed_from_u64): the PCG reference library's published demo output
//                          (pcg32_srandom_r(42, 54)); rand_core's increment per its source
//   gen_range / gen::<bool> / shuffle : rand 0.8.5 published algorithms; parity unpinned
//                          (no reference test pins the draws) beyond invariants (range,
//                          permutation, bit balance)
//   AI-REML              : dense brute-force REML score equation at convergence (numpy),
//                          IRLS KAT ai_reml.rs:489-508
//   variance ratio       : cal_cv / select KATs variance_ratio.rs:344-370, closed forms
//   score test           : KATs single_variant.rs:350-402, closed forms, allele flip
//   SPA                  : KATs binary.rs:246-307; K', K'' against finite differences of K
//   chi2 / normal tails  : statrs 0.18.0 is third-party; restated as erf/erfc closed forms
//                          (parity unpinned at the last bit)
//   BGEN layout-2 decode : the reference's .bgen fixtures against the genotype calls of the
//                          same markers in their .vcf files (tests/golden/bgen_*)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>

namespace {

// ---------------------------------------------------------------------------------
// rand_core 0.6.4 / rand_chacha 0.3.1 / rand 0.8.5 replicas (third-party; published
// algorithms restated).  Call sites: ai_reml.rs:121-130, fit_null.rs:234-238,
// variance_ratio.rs:111-115.
// ---------------------------------------------------------------------------------
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                         int rounds, uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                    key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                    (uint32_t)counter, (uint32_t)(counter >> 32),
                    (uint32_t)stream, (uint32_t)(stream >> 32)};
  uint32_t x[16];
  memcpy(x, s, sizeof(x));
#define QR(a, b, c, d)                                   \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);          \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);          \
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);           \
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
  for (int r = 0; r < rounds; r += 2) {
    QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
    QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
  }
#undef QR
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

// PCG32 (O'Neill's pcg32_random_r): 64-bit LCG state, XSH-RR output of the state before
// each step.
static const uint64_t kPcgMul = 6364136223846793005ull;
static inline uint32_t pcg32_output(uint64_t state) {
  uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
  uint32_t rot = (uint32_t)(state >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}
static void pcg32_stream(uint64_t state, uint64_t inc, size_t count, uint32_t* out) {
  for (size_t i = 0; i < count; ++i) {
    out[i] = pcg32_output(state);
    state = state * kPcgMul + inc;
  }
}

// rand_core 0.6.4 SeedableRng::seed_from_u64: PCG32 with increment INC fills the 32-byte
// seed, each word the output of the state after a step, i.e. the PCG32 stream that starts
// at seed * MUL + INC.
static void seed_from_u64(uint64_t state, uint32_t key[8]) {
  const uint64_t INC = 11634580027462260723ull;
  pcg32_stream(state * kPcgMul + INC, INC, 8, key);
}

// ChaCha8Rng (rand_chacha 0.3.1): key = seed, stream 0, 64-bit block counter from 0;
// the u32 stream is the keystream read as little-endian words.  BlockRng consumes it
// contiguously: next_u64 = word[k] | word[k+1] << 32.
struct ChaCha8Rng {
  uint32_t key[8];
  uint64_t block = 0;
  uint32_t buf[16];
  int idx = 16;
  explicit ChaCha8Rng(uint64_t seed) { seed_from_u64(seed, key); }
  uint32_t next_u32() {
    if (idx >= 16) { chacha_block(key, block++, 0, 8, buf); idx = 0; }
    return buf[idx++];
  }
  uint64_t next_u64() {
    uint64_t lo = next_u32();
    uint64_t hi = next_u32();
    return (hi << 32) | lo;
  }
  // rand 0.8.5 Standard for bool: (next_u32() as i32) < 0.
  bool gen_bool() { return (int32_t)next_u32() < 0; }
  // rand 0.8.5 UniformInt<usize>::sample_single_inclusive(0, n-1) (64-bit usize).
  uint64_t gen_range_u64(uint64_t n) {
    uint64_t range = n;
    uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
      uint64_t v = next_u64();
      unsigned __int128 m = (unsigned __int128)v * range;
      uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
      if (lo <= zone) return hi;
    }
  }
  // rand 0.8.5 UniformInt<u32>::sample_single_inclusive(0, n-1).
  uint32_t gen_range_u32(uint32_t n) {
    uint32_t range = n;
    uint32_t zone = (range << __builtin_clz(range)) - 1;
    for (;;) {
      uint32_t v = next_u32();
      uint64_t m = (uint64_t)v * range;
      uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
      if (lo <= zone) return hi;
    }
  }
  // rand 0.8.5 SliceRandom::shuffle: for i in (1..len).rev() swap(i, gen_index(i+1)).
  void shuffle(std::vector<size_t>& v) {
    for (size_t i = v.size(); i-- > 1;) {
      size_t j = (i + 1 <= 0xffffffffull) ? (size_t)gen_range_u32((uint32_t)(i + 1))
                                          : (size_t)gen_range_u64(i + 1);
      std::swap(v[i], v[j]);
    }
  }
};

// ---------------------------------------------------------------------------------
// saige-linalg dense helpers (dense.rs) on col-major storage.
// ---------------------------------------------------------------------------------
// DenseMatrix::dot  dense.rs:220-223
static double dot(const double* a, const double* b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
// DenseMatrix::mat_vec  dense.rs:102-113 (A is nrows x ncols col-major)
static void mat_vec(const double* a, size_t nrows, size_t ncols, const double* v, double* out) {
  for (size_t i = 0; i < nrows; ++i) out[i] = 0.0;
  for (size_t j = 0; j < ncols; ++j) {
    double vj = v[j];
    for (size_t i = 0; i < nrows; ++i) out[i] += a[j * nrows + i] * vj;
  }
}
// DenseMatrix::xtwx  dense.rs:227-245
static void xtwx(const double* x, size_t n, size_t p, const double* w, double* out) {
  for (size_t j = 0; j < p; ++j)
    for (size_t k = j; k < p; ++k) {
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) s += x[j * n + i] * w[i] * x[k * n + i];
      out[k * p + j] = s;
      out[j * p + k] = s;
    }
}
// DenseMatrix::xtwv  dense.rs:249-263
static void xtwv(const double* x, size_t n, size_t p, const double* w, const double* v, double* out) {
  for (size_t j = 0; j < p; ++j) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += x[j * n + i] * w[i] * v[i];
    out[j] = s;
  }
}

// CholeskyDecomp::new / solve / inverse  decomposition.rs:35-102 (L col-major p x p)
static bool cholesky(const double* a, size_t n, double* l) {
  for (size_t i = 0; i < n * n; ++i) l[i] = 0.0;
  for (size_t j = 0; j < n; ++j) {
    double sum = 0.0;
    for (size_t k = 0; k < j; ++k) sum += l[k * n + j] * l[k * n + j];
    double diag = a[j * n + j] - sum;
    if (diag <= 0.0) return false;
    l[j * n + j] = std::sqrt(diag);
    for (size_t i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (size_t k = 0; k < j; ++k) s += l[k * n + i] * l[k * n + j];
      l[j * n + i] = (a[j * n + i] - s) / l[j * n + j];
    }
  }
  return true;
}
static void chol_solve(const double* l, size_t n, const double* b, double* x) {
  std::vector<double> y(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    double sum = 0.0;
    for (size_t j = 0; j < i; ++j) sum += l[j * n + i] * y[j];
    y[i] = (b[i] - sum) / l[i * n + i];
  }
  for (size_t ii = n; ii-- > 0;) {
    double sum = 0.0;
    for (size_t j = ii + 1; j < n; ++j) sum += l[ii * n + j] * x[j];
    x[ii] = (y[ii] - sum) / l[ii * n + ii];
  }
}
static void chol_inverse(const double* l, size_t n, double* inv) {
  std::vector<double> e(n), col(n);
  for (size_t j = 0; j < n; ++j) {
    std::fill(e.begin(), e.end(), 0.0);
    e[j] = 1.0;
    chol_solve(l, n, e.data(), col.data());
    for (size_t i = 0; i < n; ++i) inv[j * n + i] = col[i];
  }
}

// ---------------------------------------------------------------------------------
// OnTheFlyGrm  saige-core/src/glmm/pcg.rs:91-181
// ---------------------------------------------------------------------------------
struct Grm {
  size_t n = 0, m = 0;
  int mode = 0;  // 0 = f32 (saige-rs), 1 = f64
  std::vector<float> z32;
  std::vector<double> z64;
  bool identity = false;
};

// OnTheFlyGrm::new  pcg.rs:101-122
static Grm* grm_new(const double* dosages, size_t n, size_t m, const double* afs, int mode) {
  Grm* g = new Grm();
  g->n = n; g->m = m; g->mode = mode;
  if (mode == 0) g->z32.assign(n * m, 0.0f); else g->z64.assign(n * m, 0.0);
  for (size_t j = 0; j < m; ++j) {
    double af = afs[j];
    double denom = std::sqrt(2.0 * af * (1.0 - af));
    if (denom > 1e-10) {
      double mean = 2.0 * af;
      for (size_t i = 0; i < n; ++i) {
        double gv = dosages[j * n + i];
        double gg = std::isnan(gv) ? mean : gv;
        double z = (gg - mean) / denom;
        if (mode == 0) g->z32[j * n + i] = (float)z; else g->z64[j * n + i] = z;
      }
    }
  }
  return g;
}

// OnTheFlyGrm::mat_vec  pcg.rs:128-180 (one rayon thread: one chunk, sequential order)
static void grm_mat_vec(const Grm* g, const double* v, double* out) {
  size_t n = g->n;
  if (g->identity) { for (size_t i = 0; i < n; ++i) out[i] = v[i]; return; }
  if (g->m == 0) { for (size_t i = 0; i < n; ++i) out[i] = 0.0; return; }
  if (g->mode == 0) {
    std::vector<float> vf(n), acc(n, 0.0f);
    for (size_t i = 0; i < n; ++i) vf[i] = (float)v[i];
    float scale = 1.0f / (float)g->m;
    for (size_t j = 0; j < g->m; ++j) {
      const float* col = &g->z32[j * n];
      float d = 0.0f;
      for (size_t i = 0; i < n; ++i) d += col[i] * vf[i];
      for (size_t i = 0; i < n; ++i) acc[i] += col[i] * d;
    }
    for (size_t i = 0; i < n; ++i) out[i] = (double)(acc[i] * scale);
  } else {
    std::vector<double> acc(n, 0.0);
    double scale = 1.0 / (double)g->m;
    for (size_t j = 0; j < g->m; ++j) {
      const double* col = &g->z64[j * n];
      double d = 0.0;
      for (size_t i = 0; i < n; ++i) d += col[i] * v[i];
      for (size_t i = 0; i < n; ++i) acc[i] += col[i] * d;
    }
    for (size_t i = 0; i < n; ++i) out[i] = acc[i] * scale;
  }
}

// ---------------------------------------------------------------------------------
// PcgSolver::solve  saige-linalg/src/decomposition.rs:212-298, with the SAIGE Sigma
// operator and diagonal preconditioner closures of ai_reml.rs:142-164 (identical in
// fit_null.rs:320-341).
// ---------------------------------------------------------------------------------
struct SigmaOp {
  const Grm* grm;
  const double* w;
  double tau0, tau1;
  size_t n;
  long long* products;
};
static void sigma_apply(const SigmaOp& s, const double* v, double* out, std::vector<double>& tmp) {
  tmp.resize(s.n);
  grm_mat_vec(s.grm, v, tmp.data());
  if (s.products) ++*s.products;
  for (size_t i = 0; i < s.n; ++i) out[i] = s.tau0 * v[i] / std::max(s.w[i], 1e-30) + s.tau1 * tmp[i];
}
static void precond_apply(const SigmaOp& s, const double* v, double* out) {
  for (size_t i = 0; i < s.n; ++i) {
    double diag = s.tau0 / std::max(s.w[i], 1e-30) + s.tau1;
    out[i] = (std::fabs(diag) > 1e-30) ? v[i] / diag : v[i];
  }
}

struct PcgOut { int iterations; double residual; bool converged; };

template <class MatVec, class Precond>
static PcgOut pcg_solve(MatVec mv, Precond pc, const double* b, size_t n, double tol, int max_iter,
                        double* x) {
  std::vector<double> r(n), z(n), p(n), ap(n);
  for (size_t i = 0; i < n; ++i) x[i] = 0.0;
  mv(x, ap.data());                      // A * x0 (x0 = 0), decomposition.rs:226
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - ap[i];
  pc(r.data(), z.data());
  p = z;
  double rz = dot(r.data(), z.data(), n);
  double b_norm = 0.0;
  for (size_t i = 0; i < n; ++i) b_norm += b[i] * b[i];
  b_norm = std::sqrt(b_norm);
  double tol_abs = tol * std::max(b_norm, 1.0);
  for (int iter = 0; iter < max_iter; ++iter) {
    double r_norm = 0.0;
    for (size_t i = 0; i < n; ++i) r_norm += r[i] * r[i];
    r_norm = std::sqrt(r_norm);
    if (r_norm < tol_abs) return {iter, r_norm, true};
    mv(p.data(), ap.data());
    double pap = dot(p.data(), ap.data(), n);
    if (std::fabs(pap) < 1e-30) return {iter, r_norm, false};
    double alpha = rz / pap;
    for (size_t i = 0; i < n; ++i) x[i] += alpha * p[i];
    for (size_t i = 0; i < n; ++i) r[i] -= alpha * ap[i];
    pc(r.data(), z.data());
    double rz_new = dot(r.data(), z.data(), n);
    double beta = rz_new / rz;
    rz = rz_new;
    for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  double r_norm = 0.0;
  for (size_t i = 0; i < n; ++i) r_norm += r[i] * r[i];
  r_norm = std::sqrt(r_norm);
  return {max_iter, r_norm, r_norm < tol_abs};
}

static PcgOut sigma_solve(const SigmaOp& s, const double* b, double tol, int max_iter, double* x) {
  std::vector<double> tmp;
  auto mv = [&](const double* v, double* out) { sigma_apply(s, v, out, tmp); };
  auto pc = [&](const double* v, double* out) { precond_apply(s, v, out); };
  return pcg_solve(mv, pc, b, s.n, tol, max_iter, x);
}

// ---------------------------------------------------------------------------------
// Family / LogitLink  family.rs:15-91, link.rs:27-46 (Binary=0, Quantitative=1)
// ---------------------------------------------------------------------------------
static inline double inv_logit(double e) { return 1.0 / (1.0 + std::exp(-e)); }
static void update_mu(int trait, const double* eta, size_t n, double* mu) {
  for (size_t i = 0; i < n; ++i) {
    if (trait == 0) {
      double m = inv_logit(eta[i]);
      mu[i] = std::min(std::max(m, 1e-10), 1.0 - 1e-10);
    } else {
      mu[i] = eta[i];
    }
  }
}
static void working_weights(int trait, const double* mu, size_t n, double* w) {
  for (size_t i = 0; i < n; ++i) {
    double v = (trait == 0) ? mu[i] * (1.0 - mu[i]) : 1.0;
    w[i] = (v > 1e-30) ? v : 1e-30;
  }
}

// fit_glm_irls  ai_reml.rs:431-482
static bool fit_glm_irls(const double* y, const double* x, size_t n, size_t p, int trait,
                         int max_iter, double* alpha) {
  std::vector<double> mu(n), eta(n), w(n), z(n), xw(p * p), xz(p), l(p * p), an(p);
  for (size_t i = 0; i < n; ++i) mu[i] = (trait == 0) ? (y[i] + 0.5) / 2.0 : y[i];
  for (size_t i = 0; i < n; ++i) eta[i] = (trait == 0) ? std::log(mu[i] / (1.0 - mu[i])) : mu[i];
  for (size_t j = 0; j < p; ++j) alpha[j] = 0.0;
  for (int it = 0; it < max_iter; ++it) {
    working_weights(trait, mu.data(), n, w.data());
    for (size_t i = 0; i < n; ++i) z[i] = eta[i] + (y[i] - mu[i]) / std::max(w[i], 1e-30);
    xtwx(x, n, p, w.data(), xw.data());
    xtwv(x, n, p, w.data(), z.data(), xz.data());
    if (!cholesky(xw.data(), p, l.data())) {
      for (size_t i = 0; i < p; ++i) xw[i * p + i] += 1e-6;
      if (!cholesky(xw.data(), p, l.data())) return false;
    }
    chol_solve(l.data(), p, xz.data(), an.data());
    mat_vec(x, n, p, an.data(), eta.data());
    update_mu(trait, eta.data(), n, mu.data());
    double change = 0.0;
    for (size_t j = 0; j < p; ++j) change += std::fabs(an[j] - alpha[j]);
    for (size_t j = 0; j < p; ++j) alpha[j] = an[j];
    if (change < 1e-8) break;
  }
  return true;
}

// ---------------------------------------------------------------------------------
// statrs 0.18.0 CDFs at the call sites single_variant.rs:179-180, binary.rs:162-177,
// restated as closed forms (third-party; last-bit parity unpinned).
// ---------------------------------------------------------------------------------
static inline double chisq1_cdf(double t) { return (t <= 0.0) ? 0.0 : std::erf(std::sqrt(0.5 * t)); }
static inline double norm_cdf(double z) { return 0.5 * std::erfc(-z / std::sqrt(2.0)); }

}  // namespace

// =================================================================================
// extern "C" surface used by tests/ and bench.py (cpu_baseline).
// =================================================================================
extern "C" {

// --- a1 decode: PlinkReader::decode_genotype / read_marker_raw / read_marker
//     saige-geno/src/plink.rs:170-179, 182-194, 237-267
void orc_decode_marker(const uint8_t* marker_bytes, size_t n_total, const uint32_t* sample_idx,
                       size_t n, double* out) {
  for (size_t k = 0; k < n; ++k) {
    size_t s = sample_idx ? sample_idx[k] : k;
    if (s >= n_total) { out[k] = std::numeric_limits<double>::quiet_NaN(); continue; }
    uint8_t bits = (uint8_t)((marker_bytes[s / 4] >> (2 * (s % 4))) & 3u);
    out[k] = bits == 0 ? 2.0 : bits == 1 ? std::numeric_limits<double>::quiet_NaN()
                                          : bits == 2 ? 1.0 : 0.0;
  }
}

// --- a2 MarkerData::compute_af  saige-geno/src/traits.rs:42-54
void orc_compute_af(const double* d, size_t n, double* sum_out, double* af, double* mac,
                    uint64_t* n_valid) {
  double sum = 0.0;
  size_t k = 0;
  for (size_t i = 0; i < n; ++i)
    if (!std::isnan(d[i])) { sum += d[i]; ++k; }
  *sum_out = sum;
  *af = k > 0 ? sum / (2.0 * (double)k) : 0.0;
  *mac = std::min(sum, 2.0 * (double)k - sum);
  *n_valid = k;
}

// --- RNG replicas
void orc_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                      uint32_t out[16]) {
  chacha_block(key, counter, stream, rounds, out);
}
void orc_seed_from_u64(uint64_t seed, uint32_t key[8]) { seed_from_u64(seed, key); }
void orc_pcg32_stream(uint64_t state, uint64_t inc, size_t count, uint32_t* out) {
  pcg32_stream(state, inc, count, out);
}
void orc_rng_u32_stream(uint64_t seed, size_t count, uint32_t* out) {
  ChaCha8Rng r(seed);
  for (size_t i = 0; i < count; ++i) out[i] = r.next_u32();
}
// ai_reml.rs:121-130 — probe-major: probe r, sample i consumes the (r*n+i)-th u32.
void orc_rademacher(uint64_t seed, size_t n, size_t nprobe, double* out) {
  ChaCha8Rng r(seed);
  for (size_t k = 0; k < nprobe; ++k)
    for (size_t i = 0; i < n; ++i) out[k * n + i] = r.gen_bool() ? 1.0 : -1.0;
}
// fit_null.rs:234-238
void orc_gen_range(uint64_t seed, size_t count, uint64_t n, uint64_t* out) {
  ChaCha8Rng r(seed);
  for (size_t i = 0; i < count; ++i) out[i] = r.gen_range_u64(n);
}
// variance_ratio.rs:111-115
void orc_shuffle(uint64_t seed, size_t len, uint64_t* out) {
  ChaCha8Rng r(seed);
  std::vector<size_t> v(len);
  for (size_t i = 0; i < len; ++i) v[i] = i;
  r.shuffle(v);
  for (size_t i = 0; i < len; ++i) out[i] = v[i];
}

// --- a4/a5 OnTheFlyGrm (dosages col-major n x m).  mode 0 = f32 (saige-rs), 1 = f64.
void* orc_grm_new(const double* dosages, size_t n, size_t m, const double* afs, int mode) {
  return grm_new(dosages, n, m, afs, mode);
}
void* orc_grm_identity(size_t n) {
  Grm* g = new Grm();
  g->n = n; g->identity = true;
  return g;
}
void orc_grm_free(void* h) { delete (Grm*)h; }
void orc_grm_mat_vec(void* h, const double* v, double* out) { grm_mat_vec((Grm*)h, v, out); }

// The same f32 product split the way pcg.rs:139-177 splits it over a rayon pool of
// `threads` workers: chunks of max(M / threads, 64) markers, each folded into its own f32
// accumulator, the accumulators added in chunk order.  Used to time the CPU path at the
// pool size; with threads > 1 the f32 sums are in a different order from threads = 1.
void orc_grm_mat_vec_threads(void* h, const double* v, double* out, int threads) {
  const Grm* g = (const Grm*)h;
  size_t n = g->n;
  if (g->mode != 0 || g->identity || g->m == 0 || threads <= 1) { grm_mat_vec(g, v, out); return; }
  std::vector<float> vf(n);
  for (size_t i = 0; i < n; ++i) vf[i] = (float)v[i];
  const size_t per = std::max<size_t>(g->m / (size_t)threads, 64);
  const size_t nchunks = (g->m + per - 1) / per;
  std::vector<std::vector<float>> accs(nchunks, std::vector<float>(n, 0.0f));
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (long long c = 0; c < (long long)nchunks; ++c) {
    std::vector<float>& acc = accs[(size_t)c];
    const size_t j1 = std::min(g->m, ((size_t)c + 1) * per);
    for (size_t j = (size_t)c * per; j < j1; ++j) {
      const float* col = &g->z32[j * n];
      float d = 0.0f;
      for (size_t i = 0; i < n; ++i) d += col[i] * vf[i];
      for (size_t i = 0; i < n; ++i) acc[i] += col[i] * d;
    }
  }
  float scale = 1.0f / (float)g->m;
  for (size_t c = 1; c < nchunks; ++c)
    for (size_t i = 0; i < n; ++i) accs[0][i] += accs[c][i];
  for (size_t i = 0; i < n; ++i) out[i] = (double)(accs[0][i] * scale);
}

// --- a7 PcgSolver::solve with the SAIGE Sigma closures (one RHS).
int orc_pcg_solve_sigma(void* grm, const double* w, double tau0, double tau1, const double* b,
                        size_t n, double tol, int max_iter, double* x, int* iters, double* resid) {
  SigmaOp s{(Grm*)grm, w, tau0, tau1, n, nullptr};
  PcgOut o = sigma_solve(s, b, tol, max_iter, x);
  *iters = o.iterations;
  *resid = o.residual;
  return o.converged ? 1 : 0;
}
// PcgSolver::solve_dense  decomposition.rs:301-315 (A row-major n x n)
int orc_pcg_solve_dense(const double* a, size_t n, const double* b, double tol, int max_iter,
                        double* x, int* iters, double* resid) {
  auto mv = [&](const double* v, double* out) {
    for (size_t i = 0; i < n; ++i) out[i] = 0.0;
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < n; ++i) out[i] += a[i * n + j] * v[j];
  };
  auto pc = [&](const double* v, double* out) {
    for (size_t i = 0; i < n; ++i) {
      double d = a[i * n + i];
      out[i] = std::fabs(d) > 1e-30 ? v[i] / d : v[i];
    }
  };
  PcgOut o = pcg_solve(mv, pc, b, n, tol, max_iter, x);
  *iters = o.iterations;
  *resid = o.residual;
  return o.converged ? 1 : 0;
}

// --- a9 fit_ai_reml  saige-core/src/glmm/ai_reml.rs:86-427
struct OrcRemlConfig {
  int max_iter; double tol; double pcg_tol; int pcg_max_iter; int n_random_vectors; uint64_t seed;
};
struct OrcRemlResult {
  double tau[2]; int iterations; int converged; double log_likelihood;
  long long grm_products; long long pcg_iterations;
};
// probes: R x n probe vectors (probe r at [r*n]) in place of the ChaCha8 Rademacher draws, or
// null for the reference's draws.  Test infrastructure only: sqrt(n) e_1 .. sqrt(n) e_n make the
// Hutchinson mean the exact trace, which pins the iteration against the REML optimum.
static int fit_ai_reml_impl(void* grm_h, const double* y, const double* x, size_t n, size_t p,
                            int trait, const OrcRemlConfig* cfg, OrcRemlResult* res,
                            double* mu_out, double* w_out, double* res_out, double* eta_out,
                            double* alpha_out, double* tau_trace, const double* probes) {
  const Grm* grm = (const Grm*)grm_h;
  bool fix_tau_e = (trait == 0);
  double tau[2] = {1.0, 0.0};
  std::vector<double> alpha(p), eta(n), mu(n), w(n), resid(n);
  if (!fit_glm_irls(y, x, n, p, trait, 25, alpha.data())) return -2;
  mat_vec(x, n, p, alpha.data(), eta.data());
  update_mu(trait, eta.data(), n, mu.data());

  size_t R = (size_t)cfg->n_random_vectors;
  std::vector<double> rv(R * n);
  if (probes) std::copy(probes, probes + R * n, rv.begin());
  else orc_rademacher(cfg->seed, n, R, rv.data());

  long long products = 0, pcg_iters = 0;
  bool converged = false;
  int n_iter = 0;
  std::vector<double> ytil(n), siy(n), six(p * n), xsx(p * p), xsy(p), l(p * p), py(n), gpy(n);
  std::vector<double> siu(n), grmu(n), corr(p), tvec(p), pu(n), sgpy(n);
  std::vector<double> xcol(n);
  for (int iter = 0; iter < cfg->max_iter; ++iter) {
    n_iter = iter + 1;
    working_weights(trait, mu.data(), n, w.data());
    for (size_t i = 0; i < n; ++i) resid[i] = y[i] - mu[i];
    SigmaOp s{grm, w.data(), tau[0], tau[1], n, &products};
    for (size_t i = 0; i < n; ++i) ytil[i] = eta[i] + resid[i] / std::max(w[i], 1e-30);
    PcgOut o = sigma_solve(s, ytil.data(), cfg->pcg_tol, cfg->pcg_max_iter, siy.data());
    pcg_iters += o.iterations;
    for (size_t j = 0; j < p; ++j) {
      o = sigma_solve(s, x + j * n, cfg->pcg_tol, cfg->pcg_max_iter, six.data() + j * n);
      pcg_iters += o.iterations;
    }
    for (size_t j = 0; j < p; ++j)
      for (size_t k = j; k < p; ++k) {
        double d = dot(x + k * n, six.data() + j * n, n);
        xsx[k * p + j] = d;
        xsx[j * p + k] = d;
      }
    for (size_t j = 0; j < p; ++j) xsy[j] = dot(x + j * n, siy.data(), n);
    if (!cholesky(xsx.data(), p, l.data())) break;
    chol_solve(l.data(), p, xsy.data(), alpha.data());
    mat_vec(x, n, p, alpha.data(), eta.data());
    update_mu(trait, eta.data(), n, mu.data());
    py = siy;
    for (size_t j = 0; j < p; ++j)
      for (size_t i = 0; i < n; ++i) py[i] -= six[j * n + i] * alpha[j];
    grm_mat_vec(grm, py.data(), gpy.data());
    ++products;
    double ypapy = dot(py.data(), gpy.data(), n);
    double trace = 0.0;
    for (size_t r = 0; r < R; ++r) {
      const double* u = rv.data() + r * n;
      o = sigma_solve(s, u, cfg->pcg_tol, cfg->pcg_max_iter, siu.data());
      pcg_iters += o.iterations;
      for (size_t j = 0; j < p; ++j) tvec[j] = dot(six.data() + j * n, u, n);
      chol_solve(l.data(), p, tvec.data(), corr.data());
      pu = siu;
      for (size_t j = 0; j < p; ++j)
        for (size_t i = 0; i < n; ++i) pu[i] -= six[j * n + i] * corr[j];
      grm_mat_vec(grm, u, grmu.data());
      ++products;
      trace += dot(grmu.data(), pu.data(), n);
    }
    trace /= (double)R;
    double score_g = ypapy - trace;
    o = sigma_solve(s, gpy.data(), cfg->pcg_tol, cfg->pcg_max_iter, sgpy.data());
    pcg_iters += o.iterations;
    for (size_t j = 0; j < p; ++j) tvec[j] = dot(six.data() + j * n, gpy.data(), n);
    chol_solve(l.data(), p, tvec.data(), corr.data());
    for (size_t j = 0; j < p; ++j)
      for (size_t i = 0; i < n; ++i) sgpy[i] -= six[j * n + i] * corr[j];
    double ai11 = dot(gpy.data(), sgpy.data(), n);
    double score_e = 0.0, ai00 = 0.0, ai01 = 0.0;
    if (!fix_tau_e) {
      // ai_reml.rs:279-317 (quantitative)
      double trw = 0.0;
      std::vector<double> iwu(n);
      for (size_t r = 0; r < R; ++r) {
        const double* u = rv.data() + r * n;
        o = sigma_solve(s, u, cfg->pcg_tol, cfg->pcg_max_iter, siu.data());
        pcg_iters += o.iterations;
        for (size_t j = 0; j < p; ++j) tvec[j] = dot(six.data() + j * n, u, n);
        chol_solve(l.data(), p, tvec.data(), corr.data());
        pu = siu;
        for (size_t j = 0; j < p; ++j)
          for (size_t i = 0; i < n; ++i) pu[i] -= six[j * n + i] * corr[j];
        for (size_t i = 0; i < n; ++i) iwu[i] = u[i] / std::max(w[i], 1e-30);
        trw += dot(iwu.data(), pu.data(), n);
      }
      trw /= (double)R;
      std::vector<double> iwpy(n), piwpy(n);
      for (size_t i = 0; i < n; ++i) iwpy[i] = py[i] / std::max(w[i], 1e-30);
      double ypa0py = dot(py.data(), iwpy.data(), n);
      score_e = ypa0py - trw;
      o = sigma_solve(s, iwpy.data(), cfg->pcg_tol, cfg->pcg_max_iter, piwpy.data());
      pcg_iters += o.iterations;
      for (size_t j = 0; j < p; ++j) tvec[j] = dot(six.data() + j * n, iwpy.data(), n);
      chol_solve(l.data(), p, tvec.data(), corr.data());
      for (size_t j = 0; j < p; ++j)
        for (size_t i = 0; i < n; ++i) piwpy[i] -= six[j * n + i] * corr[j];
      ai00 = dot(iwpy.data(), piwpy.data(), n);
      ai01 = dot(iwpy.data(), sgpy.data(), n);
    }
    double tau_new[2];
    double max_change;
    if (fix_tau_e) {
      if (std::fabs(ai11) < 1e-30) break;
      double d1 = score_g / ai11;
      double step = 1.0;
      double tg = tau[1] + step * d1;
      while (tg < 0.0 && step > 1e-10) { step *= 0.5; tg = tau[1] + step * d1; }
      tg = std::max(tg, 1e-10);
      max_change = std::fabs(tg - tau[1]) / (std::fabs(tg) + std::fabs(tau[1]) + cfg->tol);
      tau_new[0] = 1.0; tau_new[1] = tg;
    } else {
      double det = ai00 * ai11 - ai01 * ai01;
      if (std::fabs(det) < 1e-30) break;
      double d0 = (ai11 * score_e - ai01 * score_g) / det;
      double d1 = (-ai01 * score_e + ai00 * score_g) / det;
      double step = 1.0;
      double t0 = tau[0] + step * d0, t1 = tau[1] + step * d1;
      while ((t0 < 0.0 || t1 < 0.0) && step > 1e-10) {
        step *= 0.5; t0 = tau[0] + step * d0; t1 = tau[1] + step * d1;
      }
      t0 = std::max(t0, 1e-10); t1 = std::max(t1, 1e-10);
      double c0 = std::fabs(t0 - tau[0]) / (std::fabs(t0) + std::fabs(tau[0]) + cfg->tol);
      double c1 = std::fabs(t1 - tau[1]) / (std::fabs(t1) + std::fabs(tau[1]) + cfg->tol);
      max_change = std::max(c0, c1);
      tau_new[0] = t0; tau_new[1] = t1;
    }
    tau[0] = tau_new[0]; tau[1] = tau_new[1];
    if (tau_trace) { tau_trace[2 * iter] = tau[0]; tau_trace[2 * iter + 1] = tau[1]; }
    if (max_change < cfg->tol) { converged = true; break; }
  }
  working_weights(trait, mu.data(), n, w.data());
  for (size_t i = 0; i < n; ++i) resid[i] = y[i] - mu[i];
  double ll = 0.0;
  if (trait == 0) {
    for (size_t i = 0; i < n; ++i) {
      double mi = std::min(std::max(mu[i], 1e-15), 1.0 - 1e-15);
      ll += y[i] * std::log(mi) + (1.0 - y[i]) * std::log(1.0 - mi);
    }
  } else {
    double ss = dot(resid.data(), resid.data(), n);
    ll = -0.5 * (double)n * std::log(2.0 * M_PI * tau[0]) - 0.5 * ss / tau[0];
  }
  res->tau[0] = tau[0]; res->tau[1] = tau[1];
  res->iterations = n_iter; res->converged = converged ? 1 : 0; res->log_likelihood = ll;
  res->grm_products = products; res->pcg_iterations = pcg_iters;
  for (size_t i = 0; i < n; ++i) {
    if (mu_out) mu_out[i] = mu[i];
    if (w_out) w_out[i] = w[i];
    if (res_out) res_out[i] = resid[i];
    if (eta_out) eta_out[i] = eta[i];
  }
  if (alpha_out) for (size_t j = 0; j < p; ++j) alpha_out[j] = alpha[j];
  return 0;
}

int orc_fit_ai_reml(void* grm_h, const double* y, const double* x, size_t n, size_t p, int trait,
                    const OrcRemlConfig* cfg, OrcRemlResult* res, double* mu_out, double* w_out,
                    double* res_out, double* eta_out, double* alpha_out, double* tau_trace) {
  return fit_ai_reml_impl(grm_h, y, x, n, p, trait, cfg, res, mu_out, w_out, res_out, eta_out,
                          alpha_out, tau_trace, nullptr);
}
int orc_fit_ai_reml_probes(void* grm_h, const double* y, const double* x, size_t n, size_t p,
                           int trait, const OrcRemlConfig* cfg, OrcRemlResult* res,
                           double* mu_out, double* w_out, double* res_out, double* eta_out,
                           double* alpha_out, double* tau_trace, const double* probes) {
  return fit_ai_reml_impl(grm_h, y, x, n, p, trait, cfg, res, mu_out, w_out, res_out, eta_out,
                          alpha_out, tau_trace, probes);
}

// --- XVX^{-1} X'V  saige-cli/src/commands/fit_null.rs:289-307.
//     The reference writes element (j, i) of the p x n matrix at flat[j*n + i]
//     (fit_null.rs:301) and every consumer then reads the flat buffer as col-major
//     p x n (fit_null.rs:307, assoc_test.rs:125: element (r, c) = flat[c*p + r]).
//     For p > 1 the matrix used downstream is therefore a permutation of the intended
//     one.  The oracle restates the reference literally: this function fills the flat
//     buffer exactly as fit_null.rs does, and every consumer reads it col-major.
int orc_xvx_inv_xv(const double* x, size_t n, size_t p, const double* w, double* out) {
  std::vector<double> xvx(p * p), l(p * p), inv(p * p);
  xtwx(x, n, p, w, xvx.data());
  if (!cholesky(xvx.data(), p, l.data())) return -1;
  chol_inverse(l.data(), p, inv.data());
  for (size_t j = 0; j < p; ++j)
    for (size_t i = 0; i < n; ++i) {
      double v = 0.0;
      for (size_t k = 0; k < p; ++k) v += inv[k * p + j] * x[k * n + i] * w[i];
      out[j * n + i] = v;
    }
  return 0;
}

// --- cal_cv  variance_ratio.rs:249-261
double orc_cal_cv(const double* v, size_t n) {
  if (n <= 1) return std::numeric_limits<double>::infinity();
  double m = 0.0;
  for (size_t i = 0; i < n; ++i) m += v[i];
  m /= (double)n;
  if (std::fabs(m) < 1e-30) return std::numeric_limits<double>::infinity();
  double var = 0.0;
  for (size_t i = 0; i < n; ++i) var += (v[i] - m) * (v[i] - m);
  var /= ((double)n - 1.0);
  return (std::sqrt(var) / m) / (double)n;
}

// --- a10 estimate_variance_ratio  variance_ratio.rs:87-246 (overall VR only; the CLI
//     default use_categorical=false).  dosages col-major n x m.  Sigma^{-1} is the PCG
//     closure of fit_null.rs:320-343 (tol 1e-5, max_iter 500).
int orc_estimate_vr(void* grm_h, const double* dosages, const double* macs, size_t m,
                    const double* mu, const double* w, const double tau[2], int trait,
                    const double* x, const double* xvx_inv_xv, size_t n, size_t p,
                    int n_markers, double min_mac, double cv_cutoff, uint64_t seed,
                    double* vr_out, int* n_used, double* per_marker, long long* products_out) {
  ChaCha8Rng rng(seed);
  std::vector<size_t> idx(m);
  for (size_t i = 0; i < m; ++i) idx[i] = i;
  rng.shuffle(idx);
  long long products = 0;
  SigmaOp s{(const Grm*)grm_h, w, tau[0], tau[1], n, &products};
  std::vector<double> vr, g0(n), ag(p), xag(n), gt(n), sig(n);
  int target = n_markers;
  size_t tested = 0;
  for (size_t t = 0; t < m; ++t) {
    size_t k = idx[t];
    const double* graw = dosages + k * n;
    if (macs[k] < min_mac) continue;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += graw[i];
    double af = sum / (2.0 * (double)n);
    for (size_t i = 0; i < n; ++i) g0[i] = (af > 0.5) ? 2.0 - graw[i] : graw[i];
    double ac = 0.0;
    for (size_t i = 0; i < n; ++i) ac += g0[i];
    if (ac < 1.0) continue;
    mat_vec(xvx_inv_xv, p, n, g0.data(), ag.data());
    mat_vec(x, n, p, ag.data(), xag.data());
    for (size_t i = 0; i < n; ++i) gt[i] = g0[i] - xag[i];
    sigma_solve(s, gt.data(), 1e-5, 500, sig.data());
    double var1 = dot(gt.data(), sig.data(), n) / ac;
    double var2 = 0.0;
    for (size_t i = 0; i < n; ++i) {
      double mi = mu[i], gi = gt[i];
      if (trait == 0) var2 += mi * (1.0 - mi) * gi * gi / ac;
      else if (trait == 1) var2 += gi * gi / ac;
      else var2 += mi * gi * gi / ac;
    }
    if (var2 < 1e-30) continue;
    vr.push_back(var1 / var2);
    ++tested;
    if ((int)tested >= target) {
      double cv = orc_cal_cv(vr.data(), vr.size());
      if (cv <= cv_cutoff) break;
      target += 10;
    }
  }
  if (products_out) *products_out = products;
  if (vr.empty()) return -7;
  double mean = 0.0;
  for (double v : vr) mean += v;
  mean /= (double)vr.size();
  *vr_out = mean;
  *n_used = (int)vr.size();
  if (per_marker) for (size_t i = 0; i < vr.size(); ++i) per_marker[i] = vr[i];
  return 0;
}

// --- fit_null::run core flow  saige-cli/src/commands/fit_null.rs:199-394, from a packed
//     SNP-major .bed body (no magic; bytes_per_marker = ceil(n_total/4)), the sample
//     subset (model order), y and X (col-major n x p).  Outputs tau, VR, mu, the
//     XVX^{-1}X'V matrix and the GRM/VR marker split.
struct OrcFitNullOut {
  double tau[2]; int iterations; int converged; double vr; int vr_markers_used;
  int n_grm_markers; int n_vr_markers; long long grm_products; long long pcg_iterations;
  double log_likelihood;
};
int orc_fit_null(const uint8_t* bed, size_t bytes_per_marker, size_t n_total, size_t m_total,
                 const uint32_t* sample_idx, size_t n, const double* y, const double* x,
                 size_t p, int trait, const OrcRemlConfig* cfg, double min_maf,
                 double max_missing, int n_markers_vr, int grm_mode, OrcFitNullOut* out,
                 double* mu_out, double* xvx_inv_xv_out, uint8_t* is_grm_marker,
                 double* tau_trace, double* alpha_out, double* per_marker_vr) {
  std::vector<double> d(n);
  std::vector<size_t> passing;
  std::vector<double> pass_af, pass_mac;
  for (size_t m = 0; m < m_total; ++m) {
    orc_decode_marker(bed + m * bytes_per_marker, n_total, sample_idx, n, d.data());
    double sum, af, mac; uint64_t nv;
    orc_compute_af(d.data(), n, &sum, &af, &mac, &nv);
    double miss = 1.0 - ((double)nv / (double)n);
    if (af >= min_maf && af <= 1.0 - min_maf && miss <= max_missing) {
      passing.push_back(m); pass_af.push_back(af); pass_mac.push_back(mac);
    }
  }
  size_t np = passing.size();
  if (np == 0) return -8;
  ChaCha8Rng vr_rng(cfg->seed);
  std::vector<size_t> cand(1000);
  for (size_t i = 0; i < 1000; ++i) cand[i] = (size_t)vr_rng.gen_range_u64(np);
  std::sort(cand.begin(), cand.end());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  std::vector<uint8_t> in_vr(np, 0);
  for (size_t c : cand) if (pass_mac[c] >= 20.0) in_vr[c] = 1;
  std::vector<double> grm_dos, grm_af, vr_dos, vr_mac;
  size_t ng = 0, nv = 0;
  if (is_grm_marker) for (size_t m = 0; m < m_total; ++m) is_grm_marker[m] = 0;
  for (size_t i = 0; i < np; ++i) {
    orc_decode_marker(bed + passing[i] * bytes_per_marker, n_total, sample_idx, n, d.data());
    if (in_vr[i]) {
      vr_dos.insert(vr_dos.end(), d.begin(), d.end()); vr_mac.push_back(pass_mac[i]); ++nv;
      if (is_grm_marker) is_grm_marker[passing[i]] = 2;
    } else {
      grm_dos.insert(grm_dos.end(), d.begin(), d.end()); grm_af.push_back(pass_af[i]); ++ng;
      if (is_grm_marker) is_grm_marker[passing[i]] = 1;
    }
  }
  Grm* grm = grm_new(grm_dos.data(), n, ng, grm_af.data(), grm_mode);
  OrcRemlResult rr;
  std::vector<double> mu(n), w(n);
  int rc = orc_fit_ai_reml(grm, y, x, n, p, trait, cfg, &rr, mu.data(), w.data(), nullptr, nullptr,
                           alpha_out, tau_trace);
  if (rc != 0) { delete grm; return rc; }
  std::vector<double> a(p * n);
  if (orc_xvx_inv_xv(x, n, p, w.data(), a.data()) != 0) { delete grm; return -5; }
  double vr = 1.0; int used = 0; long long vprod = 0;
  if (nv > 0) {
    rc = orc_estimate_vr(grm, vr_dos.data(), vr_mac.data(), nv, mu.data(), w.data(), rr.tau, trait,
                         x, a.data(), n, p, n_markers_vr, 20.0, 0.001, cfg->seed, &vr, &used,
                         per_marker_vr, &vprod);
    if (rc != 0) { delete grm; return rc; }
  }
  delete grm;
  out->tau[0] = rr.tau[0]; out->tau[1] = rr.tau[1];
  out->iterations = rr.iterations; out->converged = rr.converged;
  out->vr = vr; out->vr_markers_used = used;
  out->n_grm_markers = (int)ng; out->n_vr_markers = (int)nv;
  out->grm_products = rr.grm_products + vprod; out->pcg_iterations = rr.pcg_iterations;
  out->log_likelihood = rr.log_likelihood;
  if (mu_out) for (size_t i = 0; i < n; ++i) mu_out[i] = mu[i];
  if (xvx_inv_xv_out) for (size_t i = 0; i < p * n; ++i) xvx_inv_xv_out[i] = a[i];
  return 0;
}

// --- BGEN layout 2  saige-geno/src/bgen.rs:373-452 (parse_layout2_probabilities, read_bits)
static uint64_t orc_read_bits(const uint8_t* data, size_t len, size_t byte_start,
                              size_t bit_offset, size_t n_bits) {  // bgen.rs:439-452
  uint64_t result = 0;
  for (size_t bit = 0; bit < n_bits; ++bit) {
    size_t total_bit = bit_offset + bit;
    size_t byte_idx = byte_start + total_bit / 8;
    size_t bit_idx = total_bit % 8;
    if (byte_idx < len) {
      uint64_t bit_val = (data[byte_idx] >> bit_idx) & 1u;
      result |= bit_val << bit;
    }
  }
  return result;
}

// Dosages of one decompressed block (out needs room for the block's N); returns 0, -1 for
// 0 bits per probability (bgen.rs:396-398), -2 for a block shorter than its header.
int orc_bgen_layout2_dosages(const uint8_t* data, size_t len, double* out, size_t cap,
                             size_t* n_out) {
  if (len < 8) return -2;
  uint32_t n_samples; uint16_t n_alleles;
  std::memcpy(&n_samples, data, 4);
  std::memcpy(&n_alleles, data + 4, 2);
  uint8_t min_ploidy = data[6], max_ploidy = data[7];
  if (len < 10 + (size_t)n_samples || n_samples > cap) return -2;
  std::vector<bool> missing(n_samples);
  for (size_t i = 0; i < n_samples; ++i) missing[i] = (data[8 + i] & 0x80) != 0;
  bool phased = data[8 + n_samples] != 0;
  size_t bits_per_prob = data[9 + n_samples];
  if (bits_per_prob == 0) return -1;
  uint64_t max_val = (bits_per_prob >= 64) ? ~0ull : ((1ull << bits_per_prob) - 1);
  for (size_t i = 0; i < n_samples; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
  if (n_alleles == 2 && min_ploidy == 2 && max_ploidy == 2 && !phased) {
    size_t prob_start = 10 + n_samples, bit_offset = 0;
    for (size_t i = 0; i < n_samples; ++i) {
      if (missing[i]) { bit_offset += 2 * bits_per_prob; continue; }
      double p_aa = (double)orc_read_bits(data, len, prob_start, bit_offset, bits_per_prob) /
                    (double)max_val;
      bit_offset += bits_per_prob;
      double p_ab = (double)orc_read_bits(data, len, prob_start, bit_offset, bits_per_prob) /
                    (double)max_val;
      bit_offset += bits_per_prob;
      out[i] = p_ab + 2.0 * (1.0 - p_aa - p_ab);
    }
  } else {
    for (size_t i = 0; i < n_samples; ++i)
      out[i] = missing[i] ? std::numeric_limits<double>::quiet_NaN() : 0.0;
  }
  *n_out = n_samples;
  return 0;
}

// --- a12 SPA  saige-core/src/spa/binary.rs:15-228 and spa/fast.rs:15-242
double orc_k0_binom(double t, const double* mu, const double* g, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += std::log(1.0 - mu[i] + mu[i] * std::exp(g[i] * t));
  return s;
}
double orc_k1_adj_binom(double t, const double* mu, const double* g, size_t n, double q) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double den = (1.0 - mu[i]) * std::exp(-g[i] * t) + mu[i];
    s += mu[i] * g[i] / den;
  }
  return s - q;
}
double orc_k2_binom(double t, const double* mu, const double* g, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double e = std::exp(-g[i] * t);
    double den = (1.0 - mu[i]) * e + mu[i];
    s += (1.0 - mu[i]) * mu[i] * g[i] * g[i] * e / (den * den);
  }
  return s;
}

}  // extern "C"

namespace {
struct FastSet {
  std::vector<double> g_nb, mu_nb;
  double na_mu = 0.0, na_sigma = 0.0;
};
// partition_for_fast_spa  fast.rs:185-204
static FastSet partition(const double* g, const double* mu, size_t n) {
  FastSet f;
  for (size_t i = 0; i < n; ++i) {
    if (std::fabs(g[i]) > 1e-10) { f.g_nb.push_back(g[i]); f.mu_nb.push_back(mu[i]); }
    else { f.na_mu += mu[i] * g[i]; f.na_sigma += mu[i] * (1.0 - mu[i]) * g[i] * g[i]; }
  }
  return f;
}
static double k0f(double t, const FastSet& f) {
  double s = 0.0;
  for (size_t i = 0; i < f.g_nb.size(); ++i)
    s += std::log(1.0 - f.mu_nb[i] + f.mu_nb[i] * std::exp(f.g_nb[i] * t));
  return s + f.na_mu * t + 0.5 * f.na_sigma * t * t;
}
static double k1f(double t, const FastSet& f, double q) {
  double s = 0.0;
  for (size_t i = 0; i < f.g_nb.size(); ++i) {
    double den = (1.0 - f.mu_nb[i]) * std::exp(-f.g_nb[i] * t) + f.mu_nb[i];
    s += f.mu_nb[i] * f.g_nb[i] / den;
  }
  return s + f.na_mu + f.na_sigma * t - q;
}
static double k2f(double t, const FastSet& f) {
  double s = 0.0;
  for (size_t i = 0; i < f.g_nb.size(); ++i) {
    double e = std::exp(-f.g_nb[i] * t);
    double den = (1.0 - f.mu_nb[i]) * e + f.mu_nb[i];
    s += (1.0 - f.mu_nb[i]) * f.mu_nb[i] * f.g_nb[i] * f.g_nb[i] * e / (den * den);
  }
  return s + f.na_sigma;
}
struct Root { double root; int iterations; bool converged; };
// find_root_k1  binary.rs:59-138 / find_root_k1_fast  fast.rs:53-132 (same control flow)
template <class K1, class K2>
static Root find_root(double init, double gpos, double gneg, double q, double tol, int max_iter,
                      K1 k1, K2 k2) {
  if (q >= gpos || q <= gneg) return {std::numeric_limits<double>::infinity(), 0, true};
  double t = init;
  double k1e = k1(t);
  double prev_jump = std::numeric_limits<double>::infinity();
  for (int it = 1; it <= max_iter; ++it) {
    double k2e = k2(t);
    double tn = t - k1e / k2e;
    if (std::isnan(tn) || std::isinf(tn)) return {t, it, false};
    if (std::fabs(tn - t) < tol) return {tn, it, true};
    if (it == max_iter) return {tn, it, false};
    double nk1 = k1(tn);
    if (k1e * nk1 < 0.0) {
      if (std::fabs(tn - t) > (prev_jump - tol)) {
        double dir = (nk1 - k1e > 0.0) ? 1.0 : -1.0;
        double th = t + dir * prev_jump / 2.0;
        prev_jump /= 2.0;
        t = th;
        k1e = k1(t);
      } else {
        prev_jump = std::fabs(tn - t);
        t = tn;
        k1e = nk1;
      }
    } else {
      t = tn;
      k1e = nk1;
    }
  }
  return {t, max_iter, false};
}
// saddle_probability  binary.rs:143-180 / fast.rs:135-180 (log_p = false)
template <class K0, class K2>
static double saddle_prob(double zeta, double q, bool* ok, K0 k0, K2 k2) {
  double k1 = k0(zeta);
  double kk2 = k2(zeta);
  double temp1 = zeta * q - k1;
  if (!std::isfinite(k1) || !std::isfinite(kk2) || temp1 < 0.0 || kk2 < 0.0) { *ok = false; return 0.0; }
  double sgn = (zeta > 0.0) ? 1.0 : (zeta < 0.0 ? -1.0 : (std::signbit(zeta) ? -1.0 : 1.0));
  if (std::isnan(zeta)) sgn = zeta;
  double w = sgn * std::sqrt(2.0 * temp1);
  double v = zeta * std::sqrt(kk2);
  if (w == 0.0) { *ok = false; return 0.0; }
  double z = w + (1.0 / w) * std::log(v / w);
  *ok = true;
  if (z > 0.0) return 1.0 - norm_cdf(z);
  return -norm_cdf(z);
}
}  // namespace

extern "C" {

struct OrcSpaResult { double pvalue; int converged; int is_spa; int iters1; int iters2;
                      double root1; double root2; };

// spa_binary  binary.rs:194-228 (fast = 0) and spa_binary_fast  fast.rs:207-242 (fast = 1)
void orc_spa_binary(const double* mu, const double* g, size_t n, double q, double qinv,
                    double pval_noadj, double tol, int fast, double init1, double init2,
                    OrcSpaResult* out) {
  double gpos = 0.0, gneg = 0.0;
  for (size_t i = 0; i < n; ++i) { if (g[i] > 0.0) gpos += g[i]; }
  for (size_t i = 0; i < n; ++i) { if (g[i] < 0.0) gneg += g[i]; }
  Root r1, r2;
  bool ok1 = false, ok2 = false;
  double p1 = 0.0, p2 = 0.0;
  if (fast) {
    FastSet f = partition(g, mu, n);
    auto k1a = [&](double t) { return k1f(t, f, q); };
    auto k1b = [&](double t) { return k1f(t, f, qinv); };
    auto k2 = [&](double t) { return k2f(t, f); };
    auto k0 = [&](double t) { return k0f(t, f); };
    r1 = find_root(init1, gpos, gneg, q, tol, 1000, k1a, k2);
    r2 = find_root(init2, gpos, gneg, qinv, tol, 1000, k1b, k2);
    if (r1.converged && r2.converged) {
      p1 = saddle_prob(r1.root, q, &ok1, k0, k2);
      p2 = saddle_prob(r2.root, qinv, &ok2, k0, k2);
    }
  } else {
    auto k1a = [&](double t) { return orc_k1_adj_binom(t, mu, g, n, q); };
    auto k1b = [&](double t) { return orc_k1_adj_binom(t, mu, g, n, qinv); };
    auto k2 = [&](double t) { return orc_k2_binom(t, mu, g, n); };
    auto k0 = [&](double t) { return orc_k0_binom(t, mu, g, n); };
    r1 = find_root(init1, gpos, gneg, q, tol, 1000, k1a, k2);
    r2 = find_root(init2, gpos, gneg, qinv, tol, 1000, k1b, k2);
    if (r1.converged && r2.converged) {
      p1 = saddle_prob(r1.root, q, &ok1, k0, k2);
      p2 = saddle_prob(r2.root, qinv, &ok2, k0, k2);
    }
  }
  out->iters1 = r1.iterations; out->iters2 = r2.iterations;
  out->root1 = r1.root; out->root2 = r2.root;
  if (r1.converged && r2.converged) {
    double a = ok1 ? p1 : pval_noadj / 2.0;
    double b = ok2 ? p2 : pval_noadj / 2.0;
    out->pvalue = std::fabs(a) + std::fabs(b);
    out->converged = 1;
    out->is_spa = (ok1 || ok2) ? 1 : 0;
  } else {
    out->pvalue = pval_noadj;
    out->converged = 0;
    out->is_spa = 0;
  }
}

// --- a11 ScoreTestEngine::test_marker  single_variant.rs:112-257 (+ compute_case_control_af
//     :260-291).  X col-major n x p; A = XVX^{-1}X'V col-major p x n.
struct OrcEngine {
  const double* mu; const double* mu2; const double* res; double tau_e; double tau_g;
  const double* a; const double* x; size_t n; size_t p; double vr;
  const double* cat_upper; const double* cat_vr; size_t n_cat;
  int use_spa; int use_fast_spa; double spa_tol; double spa_cutoff; const double* y;
};
struct OrcVariantResult {
  double ac, af, mac; uint64_t n; double score, var_t, var_t_star, tstat, pvalue, pvalue_noadj;
  int is_spa, spa_converged; double beta, se_beta, vr, af_cases, af_controls;
  uint64_t n_cases, n_controls; int spa_iters1, spa_iters2;
};
// select_variance_ratio  variance_ratio.rs:289-301
double orc_select_vr(double mac, double overall, const double* upper, const double* vr, size_t k) {
  if (k == 0) return overall;
  for (size_t i = 0; i < k; ++i) if (mac <= upper[i]) return vr[i];
  return overall;
}
void orc_test_marker(const OrcEngine* e, const double* g, OrcVariantResult* r) {
  size_t n = e->n, p = e->p;
  double sum = 0.0; size_t nv = 0;
  for (size_t i = 0; i < n; ++i) if (!std::isnan(g[i])) { sum += g[i]; ++nv; }
  double af = sum / (2.0 * (double)nv);
  double ac = sum;
  double mac = std::min(ac, 2.0 * (double)nv - ac);
  std::vector<double> ag(p), xp(n), gt(n);
  mat_vec(e->a, p, n, g, ag.data());
  mat_vec(e->x, n, p, ag.data(), xp.data());
  for (size_t i = 0; i < n; ++i) gt[i] = g[i] - xp[i];
  double sraw = 0.0;
  for (size_t i = 0; i < n; ++i) sraw += gt[i] * e->res[i];
  double score = sraw / e->tau_e;
  double vr = orc_select_vr(mac, e->vr, e->cat_upper, e->cat_vr, e->n_cat);
  double vt = 0.0;
  for (size_t i = 0; i < n; ++i) vt += gt[i] * gt[i] * e->mu2[i];
  vt = vt * e->tau_e;
  double vts = vt * vr;
  double tstat = (vts > 1e-30) ? score * score / vts : 0.0;
  double pn = 1.0 - chisq1_cdf(tstat);
  double pval = pn; int is_spa = 0, conv = 1;
  r->spa_iters1 = 0; r->spa_iters2 = 0;
  if (e->use_spa && pn < e->spa_cutoff) {
    double m1 = 0.0;
    for (size_t i = 0; i < n; ++i) m1 += e->mu[i] * gt[i];
    double q = score / std::sqrt(vr) + m1;
    double qinv = 2.0 * m1 - q;
    OrcSpaResult s;
    orc_spa_binary(e->mu, gt.data(), n, q, qinv, pn, e->spa_tol, e->use_fast_spa, 0.0, 0.0, &s);
    pval = s.pvalue; is_spa = s.is_spa; conv = s.converged;
    r->spa_iters1 = s.iters1; r->spa_iters2 = s.iters2;
  }
  r->beta = (vts > 1e-30) ? score / vts : 0.0;
  r->se_beta = (vts > 1e-30) ? 1.0 / std::sqrt(vts) : std::numeric_limits<double>::infinity();
  if (e->y) {
    double sc = 0.0, sn = 0.0; size_t nc = 0, nn = 0;
    for (size_t i = 0; i < n; ++i) {
      if (std::isnan(g[i])) continue;
      if (e->y[i] > 0.5) { sc += g[i]; ++nc; } else { sn += g[i]; ++nn; }
    }
    r->af_cases = nc > 0 ? sc / (2.0 * (double)nc) : std::numeric_limits<double>::quiet_NaN();
    r->af_controls = nn > 0 ? sn / (2.0 * (double)nn) : std::numeric_limits<double>::quiet_NaN();
    r->n_cases = nc; r->n_controls = nn;
  } else {
    r->af_cases = r->af_controls = std::numeric_limits<double>::quiet_NaN();
    r->n_cases = r->n_controls = 0;
  }
  r->ac = ac; r->af = af; r->mac = mac; r->n = nv; r->score = score; r->var_t = vt;
  r->var_t_star = vts; r->tstat = tstat; r->pvalue = pval; r->pvalue_noadj = pn;
  r->is_spa = is_spa; r->spa_converged = conv; r->vr = vr;
}

// Step-2 over a variant-major batch of uint8 dosages (dosage = q * scale), one variant
// after another exactly as assoc_test.rs:165-215 loops (single thread).
void orc_test_batch_u8(const OrcEngine* e, const uint8_t* q, size_t ld, size_t nvar, double scale,
                       OrcVariantResult* out) {
  std::vector<double> g(e->n);
  for (size_t v = 0; v < nvar; ++v) {
    for (size_t i = 0; i < e->n; ++i) g[i] = (double)q[v * ld + i] * scale;
    orc_test_marker(e, g.data(), out + v);
  }
}

// Plain chi-square(1) and normal CDFs as used at the statrs call sites.
double orc_chisq1_sf(double t) { return 1.0 - chisq1_cdf(t); }
double orc_norm_cdf(double z) { return norm_cdf(z); }

}  // extern "C"
