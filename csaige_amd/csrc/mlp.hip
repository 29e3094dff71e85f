// Training and inference of the learned saddle-point root proposal (mlp_common.hpp).
//
// Inference uses bf16-rounded weights with fp32 accumulation.  Training keeps fp32
// master weights: one work-group of 64 threads handles 64 samples (one per thread) for
// the forward and backward passes, keeping activations and deltas in LDS, then the 64
// threads turn the 64 samples into a partial gradient of every parameter; a second
// kernel sums the partials of all work-groups in a fixed order and applies Adam.  The
// gradients are therefore deterministic.
#include "common.hpp"
#include "mlp_common.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace saige {

struct MlpNorm {
  float mean[kMlpIn];
  float inv_std[kMlpIn];
  float y_scale;
};

// forward pass of one sample with weights w (fp32), input xn (standardized)
__device__ __forceinline__ float mlp_forward_one(const float* __restrict__ w, const float* xn) {
  float h1[kMlpH];
#pragma unroll 4
  for (int j = 0; j < kMlpH; ++j) {
    float a = w[kMlpB1 + j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) a = fmaf(w[kMlpW1 + j * kMlpIn + i], xn[i], a);
    h1[j] = mlp_gelu(a);
  }
  float out = w[kMlpB3];
  for (int j = 0; j < kMlpH; ++j) {
    float a = w[kMlpB2 + j];
#pragma unroll 8
    for (int k = 0; k < kMlpH; ++k) a = fmaf(w[kMlpW2 + j * kMlpH + k], h1[k], a);
    out = fmaf(w[kMlpW3 + j], mlp_gelu(a), out);
  }
  return out;
}

__global__ void mlp_predict_kernel(const float* __restrict__ w, MlpNorm nm,
                                   const float* __restrict__ x, size_t n, float* __restrict__ y) {
  __shared__ float sw[kMlpParams];
  for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) sw[p] = w[p];
  __syncthreads();
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xn[kMlpIn];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) xn[i] = (x[s * kMlpIn + i] - nm.mean[i]) * nm.inv_std[i];
  y[s] = sinhf(mlp_forward_one(sw, xn) * nm.y_scale);
}

constexpr int kMlpTB = 64;        // samples (threads) per training work-group
constexpr int kMlpLd = kMlpH + 1;

// Per work-group partial gradient of the mean squared error over the n samples listed in
// idx[b0 .. b0 + nb) (standardized inputs xn, standardized targets yn); scale = 1 / N.
__global__ __launch_bounds__(kMlpTB) void mlp_grad_kernel(
    const float* __restrict__ w, const float* __restrict__ xn, const float* __restrict__ yn,
    const uint32_t* __restrict__ idx, size_t b0, size_t nb, float scale,
    float* __restrict__ partial) {
  __shared__ float sw[kMlpParams];
  __shared__ float A1[kMlpTB][kMlpLd], H1[kMlpTB][kMlpLd], D2[kMlpTB][kMlpLd], H2[kMlpTB][kMlpLd];
  __shared__ float X[kMlpTB][kMlpIn];
  __shared__ float DO[kMlpTB], LS[kMlpTB];
  const int t = threadIdx.x;
  for (int p = t; p < kMlpParams; p += kMlpTB) sw[p] = w[p];
  const size_t local = (size_t)blockIdx.x * kMlpTB + t;
  const bool valid = local < nb;
  float y = 0.0f;
  if (valid) {
    const size_t s = idx ? idx[b0 + local] : b0 + local;
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) X[t][i] = xn[s * kMlpIn + i];
    y = yn[s];
  } else {
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) X[t][i] = 0.0f;
  }
  __syncthreads();
  // forward
  for (int j = 0; j < kMlpH; ++j) {
    float a = sw[kMlpB1 + j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) a = fmaf(sw[kMlpW1 + j * kMlpIn + i], X[t][i], a);
    A1[t][j] = a;
    H1[t][j] = mlp_gelu(a);
  }
  float out = sw[kMlpB3];
  for (int j = 0; j < kMlpH; ++j) {
    float a = sw[kMlpB2 + j];
    for (int k = 0; k < kMlpH; ++k) a = fmaf(sw[kMlpW2 + j * kMlpH + k], H1[t][k], a);
    const float h = mlp_gelu(a);
    H2[t][j] = h;
    D2[t][j] = mlp_gelu_grad(a);
    out = fmaf(sw[kMlpW3 + j], h, out);
  }
  const float err = valid ? out - y : 0.0f;
  const float dout = 2.0f * err * scale;
  DO[t] = dout;
  LS[t] = err * err;
  // backward: delta2 = dout * W3 * gelu'(a2); delta1 = gelu'(a1) * W2^T delta2
  for (int j = 0; j < kMlpH; ++j) D2[t][j] = dout * sw[kMlpW3 + j] * D2[t][j];
  for (int k = 0; k < kMlpH; ++k) {
    float a = 0.0f;
    for (int j = 0; j < kMlpH; ++j) a = fmaf(sw[kMlpW2 + j * kMlpH + k], D2[t][j], a);
    A1[t][k] = mlp_gelu_grad(A1[t][k]) * a;
  }
  __syncthreads();
  // partial gradients over this work-group's samples (fixed order over s)
  float* out_p = partial + (size_t)blockIdx.x * (kMlpParams + 1);
  for (int p = t; p < kMlpParams; p += kMlpTB) {
    float g = 0.0f;
    if (p < kMlpB1) {
      const int j = p / kMlpIn, i = p % kMlpIn;
      for (int s = 0; s < kMlpTB; ++s) g = fmaf(A1[s][j], X[s][i], g);
    } else if (p < kMlpW2) {
      const int j = p - kMlpB1;
      for (int s = 0; s < kMlpTB; ++s) g += A1[s][j];
    } else if (p < kMlpB2) {
      const int j = (p - kMlpW2) / kMlpH, k = (p - kMlpW2) % kMlpH;
      for (int s = 0; s < kMlpTB; ++s) g = fmaf(D2[s][j], H1[s][k], g);
    } else if (p < kMlpW3) {
      const int j = p - kMlpB2;
      for (int s = 0; s < kMlpTB; ++s) g += D2[s][j];
    } else if (p < kMlpB3) {
      const int j = p - kMlpW3;
      for (int s = 0; s < kMlpTB; ++s) g = fmaf(DO[s], H2[s][j], g);
    } else {
      for (int s = 0; s < kMlpTB; ++s) g += DO[s];
    }
    out_p[p] = g;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int s = 0; s < kMlpTB; ++s) l += LS[s];
    out_p[kMlpParams] = l;
  }
}

// Sum the partials in a fixed order; either return the gradient or take an Adam step.
__global__ void mlp_reduce_kernel(const float* __restrict__ partial, size_t nblk,
                                  float* __restrict__ w, float* __restrict__ m,
                                  float* __restrict__ v, float* __restrict__ w_inf, int step,
                                  float lr, float* __restrict__ grad_out,
                                  float* __restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float g = 0.0f;
  for (size_t b = 0; b < nblk; ++b) g += partial[b * (kMlpParams + 1) + p];
  if (p == kMlpParams) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (grad_out) {
    grad_out[p] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mm = b1 * m[p] + (1.0f - b1) * g;
  const float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.0f - powf(b1, (float)step));
  const float vh = vv / (1.0f - powf(b2, (float)step));
  const float nw = w[p] - lr * mh / (sqrtf(vh) + eps);
  w[p] = nw;
  w_inf[p] = (float)(__bf16)nw;
}

__global__ void mlp_round_kernel(const float* __restrict__ w, float* __restrict__ w_inf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) w_inf[p] = (float)(__bf16)w[p];
}

__global__ void mlp_standardize_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                       size_t n, MlpNorm nm, float* __restrict__ xn,
                                       float* __restrict__ yn) {
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) xn[s * kMlpIn + i] = (x[s * kMlpIn + i] - nm.mean[i]) * nm.inv_std[i];
  if (y) yn[s] = asinhf(y[s]) / nm.y_scale;
}

}  // namespace saige

using namespace saige;

struct saige_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  MlpNorm norm{};
  DevBuf w, w_inf, m, v, partial, xn, yn, idx, x, y, out;
  int step = 0;
};

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

extern "C" {

int saige_mlp_create(int device, uint64_t seed, saige_mlp** out) {
  clear_error();
  SAIGE_ARG(out, "out");
  *out = nullptr;
  SAIGE_TRY(use_device(device));
  saige_mlp* h = new saige_mlp();
  h->device = device;
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    set_error("hipStreamCreate failed");
    return SAIGE_ERR_DEVICE;
  }
  // Glorot-uniform weights, zero biases; identity standardization
  std::vector<float> p(kMlpParams, 0.0f);
  auto uni = [&](uint64_t k) { return (float)((mix64(seed * 0x9e3779b97f4a7c15ull + k) >> 11) * (1.0 / 9007199254740992.0)); };
  const float l1 = std::sqrt(6.0f / (kMlpIn + kMlpH)), l2 = std::sqrt(6.0f / (2 * kMlpH)),
              l3 = std::sqrt(6.0f / (kMlpH + 1));
  for (int q = 0; q < kMlpH * kMlpIn; ++q) p[kMlpW1 + q] = (2.0f * uni(q) - 1.0f) * l1;
  for (int q = 0; q < kMlpH * kMlpH; ++q) p[kMlpW2 + q] = (2.0f * uni(10000 + q) - 1.0f) * l2;
  for (int q = 0; q < kMlpH; ++q) p[kMlpW3 + q] = (2.0f * uni(20000 + q) - 1.0f) * l3;
  for (int i = 0; i < kMlpIn; ++i) {
    h->norm.mean[i] = 0.0f;
    h->norm.inv_std[i] = 1.0f;
  }
  h->norm.y_scale = 1.0f;
  int rc = h->w.ensure(kMlpParams * 4);
  if (rc == SAIGE_OK) rc = h->w_inf.ensure(kMlpParams * 4);
  if (rc == SAIGE_OK) rc = h->m.ensure(kMlpParams * 4);
  if (rc == SAIGE_OK) rc = h->v.ensure(kMlpParams * 4);
  if (rc != SAIGE_OK) {
    saige_mlp_destroy(h);
    return rc;
  }
  if (hipMemcpy(h->w.p, p.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->m.p, 0, kMlpParams * 4) != hipSuccess ||
      hipMemset(h->v.p, 0, kMlpParams * 4) != hipSuccess) {
    saige_mlp_destroy(h);
    set_error("mlp upload failed");
    return SAIGE_ERR_DEVICE;
  }
  mlp_round_kernel<<<(kMlpParams + 255) / 256, 256, 0, h->stream>>>(h->w.as<float>(), h->w_inf.as<float>());
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    saige_mlp_destroy(h);
    set_error("mlp init failed");
    return SAIGE_ERR_DEVICE;
  }
  *out = h;
  return SAIGE_OK;
}

void saige_mlp_destroy(saige_mlp* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  delete h;
}

int saige_mlp_set_params(saige_mlp* h, const float* params, const float* norm) {
  clear_error();
  SAIGE_ARG(h && params && norm, "null pointer");
  std::lock_guard<std::mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  SAIGE_HIP(hipMemcpy(h->w.p, params, kMlpParams * 4, hipMemcpyHostToDevice));
  for (int i = 0; i < kMlpIn; ++i) {
    h->norm.mean[i] = norm[i];
    h->norm.inv_std[i] = norm[kMlpIn + i];
  }
  h->norm.y_scale = norm[2 * kMlpIn];
  mlp_round_kernel<<<(kMlpParams + 255) / 256, 256, 0, h->stream>>>(h->w.as<float>(), h->w_inf.as<float>());
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  return SAIGE_OK;
}

int saige_mlp_get_params(saige_mlp* h, float* params, float* norm) {
  clear_error();
  SAIGE_ARG(h, "null pointer");
  std::lock_guard<std::mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  if (params) SAIGE_HIP(hipMemcpy(params, h->w.p, kMlpParams * 4, hipMemcpyDeviceToHost));
  if (norm) {
    for (int i = 0; i < kMlpIn; ++i) {
      norm[i] = h->norm.mean[i];
      norm[kMlpIn + i] = h->norm.inv_std[i];
    }
    norm[2 * kMlpIn] = h->norm.y_scale;
  }
  return SAIGE_OK;
}

int saige_mlp_predict(saige_mlp* h, const float* x, size_t n, float* y) {
  clear_error();
  SAIGE_ARG(h && (n == 0 || (x && y)), "null pointer");
  if (n == 0) return SAIGE_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  SAIGE_TRY(h->x.ensure(n * kMlpIn * 4));
  SAIGE_TRY(h->out.ensure(n * 4));
  SAIGE_HIP(hipMemcpyAsync(h->x.p, x, n * kMlpIn * 4, hipMemcpyHostToDevice, h->stream));
  mlp_predict_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, h->stream>>>(h->w_inf.as<float>(), h->norm,
                                                                      h->x.as<float>(), n,
                                                                      h->out.as<float>());
  SAIGE_HIP(hipGetLastError());
  SAIGE_HIP(hipMemcpyAsync(y, h->out.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  return SAIGE_OK;
}

// Mean squared error (in standardized target units) and its gradient with respect to the
// fp32 master parameters, over n raw samples (standardized with the model's norm).
int saige_mlp_loss_grad(saige_mlp* h, const float* x, const float* y, size_t n, float* loss,
                        float* grad) {
  clear_error();
  SAIGE_ARG(h && x && y && loss && grad && n > 0, "null pointer");
  std::lock_guard<std::mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const size_t nblk = ceil_div(n, kMlpTB);
  SAIGE_TRY(h->x.ensure(n * kMlpIn * 4));
  SAIGE_TRY(h->y.ensure(n * 4));
  SAIGE_TRY(h->xn.ensure(n * kMlpIn * 4));
  SAIGE_TRY(h->yn.ensure(n * 4));
  SAIGE_TRY(h->partial.ensure(nblk * (kMlpParams + 1) * 4));
  SAIGE_TRY(h->out.ensure((kMlpParams + 1) * 4));
  SAIGE_HIP(hipMemcpyAsync(h->x.p, x, n * kMlpIn * 4, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(h->y.p, y, n * 4, hipMemcpyHostToDevice, s));
  mlp_standardize_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(
      h->x.as<float>(), h->y.as<float>(), n, h->norm, h->xn.as<float>(), h->yn.as<float>());
  mlp_grad_kernel<<<(unsigned)nblk, kMlpTB, 0, s>>>(h->w.as<float>(), h->xn.as<float>(),
                                                     h->yn.as<float>(), nullptr, 0, n,
                                                     1.0f / (float)n, h->partial.as<float>());
  float* g = h->out.as<float>();
  mlp_reduce_kernel<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(
      h->partial.as<float>(), nblk, nullptr, nullptr, nullptr, nullptr, 0, 0.0f, g, g + kMlpParams);
  SAIGE_HIP(hipGetLastError());
  std::vector<float> hg(kMlpParams + 1);
  SAIGE_HIP(hipMemcpyAsync(hg.data(), g, (kMlpParams + 1) * 4, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  std::copy(hg.begin(), hg.begin() + kMlpParams, grad);
  *loss = hg[kMlpParams] / (float)n;
  return SAIGE_OK;
}

// Adam on the mean squared error; standardization is fitted to (x, y) first.  The epoch
// order comes from a seeded shuffle.  loss_history[e] = mean loss over epoch e.
int saige_mlp_train(saige_mlp* h, const float* x, const float* y, size_t n, uint32_t epochs,
                    uint32_t batch, float lr, uint64_t seed, float* loss_history) {
  clear_error();
  SAIGE_ARG(h && x && y && n > 0 && batch > 0, "null pointer");
  std::lock_guard<std::mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  // standardization from the training data (fp64 host sums)
  for (int i = 0; i < kMlpIn; ++i) {
    double mu = 0.0, sq = 0.0;
    for (size_t k = 0; k < n; ++k) mu += x[k * kMlpIn + i];
    mu /= (double)n;
    for (size_t k = 0; k < n; ++k) sq += (x[k * kMlpIn + i] - mu) * (x[k * kMlpIn + i] - mu);
    double sd = std::sqrt(sq / (double)n);
    h->norm.mean[i] = (float)mu;
    h->norm.inv_std[i] = (float)(sd > 1e-12 ? 1.0 / sd : 1.0);
  }
  {
    double sq = 0.0;
    for (size_t k = 0; k < n; ++k) sq += std::asinh((double)y[k]) * std::asinh((double)y[k]);
    double sc = std::sqrt(sq / (double)n);
    h->norm.y_scale = (float)(sc > 1e-12 ? sc : 1.0);
  }
  const size_t nb_max = ceil_div(batch, kMlpTB);
  SAIGE_TRY(h->x.ensure(n * kMlpIn * 4));
  SAIGE_TRY(h->y.ensure(n * 4));
  SAIGE_TRY(h->xn.ensure(n * kMlpIn * 4));
  SAIGE_TRY(h->yn.ensure(n * 4));
  SAIGE_TRY(h->idx.ensure(n * 4));
  SAIGE_TRY(h->partial.ensure(nb_max * (kMlpParams + 1) * 4));
  SAIGE_TRY(h->out.ensure(64));
  SAIGE_HIP(hipMemcpyAsync(h->x.p, x, n * kMlpIn * 4, hipMemcpyHostToDevice, s));
  SAIGE_HIP(hipMemcpyAsync(h->y.p, y, n * 4, hipMemcpyHostToDevice, s));
  mlp_standardize_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(
      h->x.as<float>(), h->y.as<float>(), n, h->norm, h->xn.as<float>(), h->yn.as<float>());
  std::vector<uint32_t> perm(n);
  std::vector<float> eloss(epochs, 0.0f);
  uint64_t st = seed;
  for (uint32_t e = 0; e < epochs; ++e) {
    std::iota(perm.begin(), perm.end(), 0u);
    for (size_t i = n; i > 1; --i) {
      st = mix64(st + 0x9e3779b97f4a7c15ull);
      std::swap(perm[i - 1], perm[st % i]);
    }
    SAIGE_HIP(hipMemcpyAsync(h->idx.p, perm.data(), n * 4, hipMemcpyHostToDevice, s));
    double lsum = 0.0;
    std::vector<float> lbuf;
    size_t nsteps = ceil_div(n, batch);
    std::vector<float> losses(nsteps);
    float* dloss = h->out.as<float>();
    for (size_t b0 = 0, k = 0; b0 < n; b0 += batch, ++k) {
      const size_t nb = std::min<size_t>(batch, n - b0);
      const size_t nblk = ceil_div(nb, kMlpTB);
      mlp_grad_kernel<<<(unsigned)nblk, kMlpTB, 0, s>>>(h->w.as<float>(), h->xn.as<float>(),
                                                         h->yn.as<float>(), h->idx.as<uint32_t>(),
                                                         b0, nb, 1.0f / (float)nb,
                                                         h->partial.as<float>());
      ++h->step;
      mlp_reduce_kernel<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(
          h->partial.as<float>(), nblk, h->w.as<float>(), h->m.as<float>(), h->v.as<float>(),
          h->w_inf.as<float>(), h->step, lr, nullptr, dloss);
      SAIGE_HIP(hipGetLastError());
      float l = 0.0f;
      SAIGE_HIP(hipMemcpyAsync(&l, dloss, 4, hipMemcpyDeviceToHost, s));
      SAIGE_HIP(hipStreamSynchronize(s));
      lsum += l;
    }
    eloss[e] = (float)(lsum / (double)n);
  }
  if (loss_history) std::copy(eloss.begin(), eloss.end(), loss_history);
  return SAIGE_OK;
}

}  // extern "C"

namespace saige {
// View of a trained model for the SPA kernel (bf16-rounded weights, fp32 arithmetic).
void mlp_view(const saige_mlp* h, const float** w, float* mean, float* inv_std, float* y_scale) {
  *w = h->w_inf.as<float>();
  for (int i = 0; i < kMlpIn; ++i) {
    mean[i] = h->norm.mean[i];
    inv_std[i] = h->norm.inv_std[i];
  }
  *y_scale = h->norm.y_scale;
}
}  // namespace saige
