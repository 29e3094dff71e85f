// Seeded synthetic cohorts for the benchmark and parity workloads (harness code: it has
// no counterpart in saige-rs, which reads real PLINK files).
//
// Every random draw is a pure function of (seed, stream, index) through a splitmix64
// hash, so the same cohort is produced whatever the thread count, and any slice can be
// regenerated independently.  Genotypes are written as PLINK .bed bodies (SNP-major,
// ceil(n/4) bytes per marker, sample i at bits 2*(i%4) of byte i/4) with the dosage of
// allele 1 coded 2 -> 00, 1 -> 10, 0 -> 11 (the decoding of saige-geno/src/plink.rs:170-179).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
inline uint64_t h3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed * 0x9e3779b97f4a7c15ull + stream) ^ (idx + 0x632be59bd9b4e019ull));
}
inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
inline double normal(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t a = h3(seed, stream, 2 * idx), b = h3(seed, stream, 2 * idx + 1);
  double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740993.0);
  double u2 = u01(b);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
inline uint8_t code_of(int c) { return c == 2 ? 0u : (c == 1 ? 2u : 3u); }

// Polygenic scores are summed in 2^-40 fixed point: integer sums do not depend on the order
// in which the threads' partial sums meet, so a cohort is the same bit for bit on every run
// and thread count (each term is rounded once, to 9e-13).
constexpr double kPgsScale = 1099511627776.0;  // 2^40
inline int64_t pgs_fix(double v) { return (int64_t)std::llround(v * kPgsScale); }
inline double pgs_value(int64_t v) { return (double)v / kPgsScale; }

enum Stream : uint64_t {
  S_MAF = 1, S_GENO = 2, S_BETA = 3, S_COV = 4, S_NOISE = 5, S_TRANSMIT = 6, S_MAF_TEST = 7,
  S_GENO_TEST = 8, S_ETA = 9, S_Y = 10, S_DOSAGE = 11, S_DOSAGE_MAF = 12, S_DOSAGE_NOISE = 13,
  S_X2 = 14, S_CAUSAL = 15, S_PAIR = 16, S_PAIR_T = 17, S_BETA_T = 18, S_NOISE_T = 19
};

// Liability -> binary phenotype: the top `prevalence` fraction are cases.
void threshold_cases(const std::vector<double>& liab, double prevalence, double* y) {
  size_t n = liab.size();
  size_t ncase = (size_t)std::llround(prevalence * (double)n);
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return liab[a] > liab[b]; });
  for (size_t i = 0; i < n; ++i) y[i] = 0.0;
  for (size_t k = 0; k < ncase && k < n; ++k) y[order[k]] = 1.0;
}

}  // namespace

extern "C" {

// Unrelated samples under HWE (the 2,000-sample workload and its variants).
//   bed_grm: m_grm * ceil(n/4) bytes; bed_test: m_test * ceil(n/4) bytes (may be null).
//   x: n x (1 + n_cov) col-major, intercept first.  y: binary phenotype from the liability
//   PGS(h2) + cov_beta * sum(covariates) + N(0, 1 - h2), top `prevalence` are cases.
int synth_unrelated(uint64_t seed, size_t n, size_t m_grm, size_t m_test, double maf_lo,
                    double maf_hi, double h2, double prevalence, size_t n_cov, double cov_beta,
                    uint8_t* bed_grm, uint8_t* bed_test, double* maf_grm, double* maf_test,
                    double* y, double* x, double* liability_out) {
  size_t bpm = (n + 3) / 4;
  std::vector<int64_t> pgs(n, 0);
#pragma omp parallel
  {
    std::vector<int64_t> local(n, 0);
#pragma omp for schedule(static)
    for (long long jj = 0; jj < (long long)m_grm; ++jj) {
      size_t j = (size_t)jj;
      double maf = maf_lo + (maf_hi - maf_lo) * u01(h3(seed, S_MAF, j));
      maf_grm[j] = maf;
      double beta = normal(seed, S_BETA, j) * std::sqrt(h2 / (double)m_grm);
      double mean = 2.0 * maf, sd = std::sqrt(2.0 * maf * (1.0 - maf));
      uint8_t* row = bed_grm + j * bpm;
      std::memset(row, 0, bpm);
      for (size_t i = 0; i < n; ++i) {
        uint64_t h = h3(seed, S_GENO, j * n + i);
        double u1 = (double)(uint32_t)h * (1.0 / 4294967296.0);
        double u2 = (double)(uint32_t)(h >> 32) * (1.0 / 4294967296.0);
        int c = (u1 < maf) + (u2 < maf);
        row[i / 4] |= (uint8_t)(code_of(c) << (2 * (i % 4)));
        local[i] += pgs_fix(beta * ((double)c - mean) / sd);
      }
    }
#pragma omp critical
    for (size_t i = 0; i < n; ++i) pgs[i] += local[i];
  }
  if (bed_test) {
#pragma omp parallel for schedule(static)
    for (long long jj = 0; jj < (long long)m_test; ++jj) {
      size_t j = (size_t)jj;
      double maf = maf_lo + (maf_hi - maf_lo) * u01(h3(seed, S_MAF_TEST, j));
      if (maf_test) maf_test[j] = maf;
      uint8_t* row = bed_test + j * bpm;
      std::memset(row, 0, bpm);
      for (size_t i = 0; i < n; ++i) {
        uint64_t h = h3(seed, S_GENO_TEST, j * n + i);
        double u1 = (double)(uint32_t)h * (1.0 / 4294967296.0);
        double u2 = (double)(uint32_t)(h >> 32) * (1.0 / 4294967296.0);
        int c = (u1 < maf) + (u2 < maf);
        row[i / 4] |= (uint8_t)(code_of(c) << (2 * (i % 4)));
      }
    }
  }
  size_t p = 1 + n_cov;
  std::vector<double> liab(n);
  for (size_t i = 0; i < n; ++i) {
    x[i] = 1.0;
    double l = pgs_value(pgs[i]);
    for (size_t c = 0; c < n_cov; ++c) {
      double v = normal(seed, S_COV + 100 * (c + 1), i);
      x[(c + 1) * n + i] = v;
      l += cov_beta * v;
    }
    l += std::sqrt(1.0 - h2) * normal(seed, S_NOISE, i);
    liab[i] = l;
  }
  (void)p;
  threshold_cases(liab, prevalence, y);
  if (liability_out) for (size_t i = 0; i < n; ++i) liability_out[i] = liab[i];
  return 0;
}

// Nuclear families: 2 HWE founders + `kids` Mendelian offspring per family; sample order
// per family is father, mother, kid 1..kids.  n = n_fam * (2 + kids).
int synth_families(uint64_t seed, size_t n_fam, size_t kids, size_t m, double maf_lo,
                   double maf_hi, double h2, double prevalence, size_t n_cov, double cov_beta,
                   uint8_t* bed, double* maf_out, double* y, double* x) {
  size_t fs = 2 + kids;
  size_t n = n_fam * fs;
  size_t bpm = (n + 3) / 4;
  std::vector<int64_t> pgs(n, 0);
#pragma omp parallel
  {
    std::vector<int64_t> local(n, 0);
    std::vector<int> dose(fs);
#pragma omp for schedule(static)
    for (long long jj = 0; jj < (long long)m; ++jj) {
      size_t j = (size_t)jj;
      double maf = maf_lo + (maf_hi - maf_lo) * u01(h3(seed, S_MAF, j));
      maf_out[j] = maf;
      double beta = normal(seed, S_BETA, j) * std::sqrt(h2 / (double)m);
      double mean = 2.0 * maf, sd = std::sqrt(2.0 * maf * (1.0 - maf));
      uint8_t* row = bed + j * bpm;
      std::memset(row, 0, bpm);
      for (size_t f = 0; f < n_fam; ++f) {
        uint64_t h = h3(seed, S_GENO, j * n_fam + f);
        int fa0 = ((uint16_t)h) * (1.0 / 65536.0) < maf;
        int fa1 = ((uint16_t)(h >> 16)) * (1.0 / 65536.0) < maf;
        int mo0 = ((uint16_t)(h >> 32)) * (1.0 / 65536.0) < maf;
        int mo1 = ((uint16_t)(h >> 48)) * (1.0 / 65536.0) < maf;
        uint64_t t = h3(seed, S_TRANSMIT, j * n_fam + f);
        dose[0] = fa0 + fa1;
        dose[1] = mo0 + mo1;
        for (size_t k = 0; k < kids; ++k) {
          int pa = ((t >> (2 * k)) & 1) ? fa1 : fa0;
          int ma = ((t >> (2 * k + 1)) & 1) ? mo1 : mo0;
          dose[2 + k] = pa + ma;
        }
        for (size_t s = 0; s < fs; ++s) {
          size_t i = f * fs + s;
          row[i / 4] |= (uint8_t)(code_of(dose[s]) << (2 * (i % 4)));
          local[i] += pgs_fix(beta * ((double)dose[s] - mean) / sd);
        }
      }
    }
#pragma omp critical
    for (size_t i = 0; i < n; ++i) pgs[i] += local[i];
  }
  std::vector<double> liab(n);
  for (size_t i = 0; i < n; ++i) {
    x[i] = 1.0;
    double l = pgs_value(pgs[i]);
    for (size_t c = 0; c < n_cov; ++c) {
      double v = normal(seed, S_COV + 100 * (c + 1), i);
      x[(c + 1) * n + i] = v;
      l += cov_beta * v;
    }
    l += std::sqrt(1.0 - h2) * normal(seed, S_NOISE, i);
    liab[i] = l;
  }
  threshold_cases(liab, prevalence, y);
  return 0;
}

// Biobank-style cohort for several phenotypes: n samples, the first 2 * n_pairs of them in
// half-sibling pairs (2nd-degree relatives: one shared parent, two unrelated mates; pair k
// is samples 2k and 2k + 1), the rest unrelated, all under HWE with MAF ~ U(lo, hi) and no
// LD.  n_traits liabilities share the covariates but have their own causal effects
// (markers causal with probability n_causal / m, beta ~ N(0, h2 / n_causal)) and noise;
// trait t is binary with the top prevalence[t] fraction as cases.  y: n_traits x n.
int synth_biobank(uint64_t seed, size_t n, size_t m, size_t n_pairs, size_t n_traits,
                  const double* prevalence, double h2, size_t n_causal, double maf_lo,
                  double maf_hi, size_t n_cov, double cov_beta, uint8_t* bed, double* maf_out,
                  double* y, double* x) {
  if (2 * n_pairs > n || n_traits == 0) return 1;
  const size_t bpm = (n + 3) / 4;
  const size_t nrel = 2 * n_pairs;
  const double pc = std::min(1.0, (double)n_causal / (double)std::max<size_t>(m, 1));
  const double bsd = std::sqrt(h2 / (double)std::max<size_t>(n_causal, 1));
  std::vector<int64_t> pgs(n * n_traits, 0);
#pragma omp parallel
  {
    std::vector<int64_t> local;
    std::vector<uint8_t> dose(n);
#pragma omp for schedule(dynamic, 64)
    for (long long jj = 0; jj < (long long)m; ++jj) {
      const size_t j = (size_t)jj;
      const double maf = maf_lo + (maf_hi - maf_lo) * u01(h3(seed, S_MAF, j));
      maf_out[j] = maf;
      const uint32_t thr = (uint32_t)std::min(65536.0, maf * 65536.0);
      auto allele = [&](uint64_t h, int k) -> int { return ((h >> (16 * k)) & 0xFFFFu) < thr; };
      for (size_t k = 0; k < n_pairs; ++k) {
        const uint64_t a = h3(seed, S_PAIR, j * n_pairs + k);       // parent, mate 1
        const uint64_t b = h3(seed, S_PAIR + 100, j * n_pairs + k); // mate 2
        const uint64_t t = h3(seed, S_PAIR_T, j * n_pairs + k);
        const int par[2] = {allele(a, 0), allele(a, 1)};
        const int m1[2] = {allele(a, 2), allele(a, 3)};
        const int m2[2] = {allele(b, 0), allele(b, 1)};
        dose[2 * k] = (uint8_t)(par[t & 1] + m1[(t >> 1) & 1]);
        dose[2 * k + 1] = (uint8_t)(par[(t >> 2) & 1] + m2[(t >> 3) & 1]);
      }
      for (size_t i = nrel; i < n; i += 2) {
        const uint64_t h = h3(seed, S_GENO, j * n + i);
        dose[i] = (uint8_t)(allele(h, 0) + allele(h, 1));
        if (i + 1 < n) dose[i + 1] = (uint8_t)(allele(h, 2) + allele(h, 3));
      }
      uint8_t* row = bed + j * bpm;
      std::memset(row, 0, bpm);
      for (size_t i = 0; i < n; ++i) row[i / 4] |= (uint8_t)(code_of(dose[i]) << (2 * (i % 4)));
      if (u01(h3(seed, S_CAUSAL, j)) < pc) {
        if (local.empty()) local.assign(n * n_traits, 0);
        const double mean = 2.0 * maf, sd = std::sqrt(2.0 * maf * (1.0 - maf));
        for (size_t t = 0; t < n_traits; ++t) {
          const double beta = bsd * normal(seed, S_BETA_T + 100 * t, j);
          int64_t* lt = local.data() + t * n;
          for (size_t i = 0; i < n; ++i) lt[i] += pgs_fix(beta * ((double)dose[i] - mean) / sd);
        }
      }
    }
#pragma omp critical
    if (!local.empty())
      for (size_t i = 0; i < n * n_traits; ++i) pgs[i] += local[i];
  }
  std::vector<double> cov(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    x[i] = 1.0;
    for (size_t c = 0; c < n_cov; ++c) {
      const double v = normal(seed, S_COV + 100 * (c + 1), i);
      x[(c + 1) * n + i] = v;
      cov[i] += cov_beta * v;
    }
  }
  std::vector<double> liab(n);
  for (size_t t = 0; t < n_traits; ++t) {
    for (size_t i = 0; i < n; ++i)
      liab[i] = pgs_value(pgs[t * n + i]) + cov[i] +
                std::sqrt(1.0 - h2) * normal(seed, S_NOISE_T + 100 * t, i);
    threshold_cases(liab, prevalence[t], y + t * n);
  }
  return 0;
}

// Step-2 streamed workload: a fitted null with mu_i = logistic(eta_i), eta_i ~ N(eta_mean,
// eta_sd) (sd stated as a standard deviation), y_i ~ Bernoulli(mu_i); X = [1] or
// [1, N(0,1)].  Returns mu, y, x (n x p col-major).
int synth_null_fit(uint64_t seed, size_t n, size_t p, double eta_mean, double eta_sd, double* mu,
                   double* y, double* x) {
  for (size_t i = 0; i < n; ++i) {
    double eta = eta_mean + eta_sd * normal(seed, S_ETA, i);
    double m = 1.0 / (1.0 + std::exp(-eta));
    mu[i] = m;
    y[i] = (u01(h3(seed, S_Y, i)) < m) ? 1.0 : 0.0;
    x[i] = 1.0;
    for (size_t c = 1; c < p; ++c) x[c * n + i] = normal(seed, S_X2 + 100 * c, i);
  }
  return 0;
}

// MAF of step-2 variants [first, first + nvar): log-uniform(lo, hi).
int synth_dosage_mafs(uint64_t seed, size_t first, size_t nvar, double maf_lo, double maf_hi,
                      double* maf_out) {
  const double llo = std::log(maf_lo), lhi = std::log(maf_hi);
  for (size_t vv = 0; vv < nvar; ++vv)
    maf_out[vv] = std::exp(llo + (lhi - llo) * u01(h3(seed, S_DOSAGE_MAF, first + vv)));
  return 0;
}

// uint8-quantized dosages (dosage = q / 127, so 0, 1 and 2 are exact) for variants
// [first, first + nvar), variant-major rows of `ld` bytes.  MAF ~ log-uniform(lo, hi);
// hard calls under HWE; every variant with index % 4 == 3 carries imputation-like noise
// q = clamp(round(127 g + 12 e), 0, 254) with e = sqrt(3) (u1 + u2 + u3 + u4 - 2), the four
// u the 16-bit fields of one hash over 65536 (Irwin-Hall, mean 0, variance 1).  Every step is
// integer or one correctly rounded fp64 operation, so the device generator
// (dosage_gpu.hip) writes the same bytes.
int synth_dosage_u8(uint64_t seed, size_t n, size_t first, size_t nvar, size_t ld, double maf_lo,
                    double maf_hi, uint8_t* q, double* maf_out) {
  std::vector<double> mafs(nvar);
  synth_dosage_mafs(seed, first, nvar, maf_lo, maf_hi, mafs.data());
#pragma omp parallel for schedule(static)
  for (long long vv = 0; vv < (long long)nvar; ++vv) {
    size_t v = first + (size_t)vv;
    const double maf = mafs[vv];
    if (maf_out) maf_out[vv] = maf;
    uint8_t* row = q + (size_t)vv * ld;
    bool noisy = (v % 4) == 3;
    for (size_t i = 0; i < n; ++i) {
      uint64_t h = h3(seed, S_DOSAGE, v * n + i);
      double u1 = (double)(uint32_t)h * (1.0 / 4294967296.0);
      double u2 = (double)(uint32_t)(h >> 32) * (1.0 / 4294967296.0);
      int c = (u1 < maf) + (u2 < maf);
      int qq = 127 * c;
      if (noisy) {
        const uint64_t g = h3(seed, S_DOSAGE_NOISE, v * n + i);
        const uint32_t s4 = (uint32_t)(g & 0xFFFFu) + (uint32_t)((g >> 16) & 0xFFFFu) +
                            (uint32_t)((g >> 32) & 0xFFFFu) + (uint32_t)(g >> 48);
        const double t = (double)s4 * (1.0 / 65536.0) - 2.0;   // exact
        const double e = (t * 1.7320508075688772) * 12.0;
        const double y = 127.0 * (double)c + e;
        qq = (int)std::lround(y);
        qq = std::min(254, std::max(0, qq));
      }
      row[i] = (uint8_t)qq;
    }
    for (size_t i = n; i < ld; ++i) row[i] = 0;
  }
  return 0;
}

}  // extern "C"
