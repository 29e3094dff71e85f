// Learned saddle-point root proposal: a 6 -> 64 -> 64 -> 1 multilayer perceptron with GELU
// activations (no counterpart in saige-rs; the plain bracketed Newton of fast.rs:53-132
// stays the solver and the judge of the proposal).
//
// Parameter vector layout (4,673 floats): W1 (64 x 6, row j = hidden unit j), b1 (64),
// W2 (64 x 64, row j), b2 (64), W3 (64), b3 (1).
//
// Inputs per variant and root (root 2 negates the first input):
//   x0 = S / sqrt(varT*)           signed z of the score
//   x1 = log(varT)                 x2 = log(1 + MAC)
//   x3..x5 = log(1e-8 + sum over carriers (g > 0) of mu, mu(1 - mu), mu^2 (1 - mu))
// each standardized with the training mean and standard deviation.  The target is the
// saddle point in units of sqrt(K''(0)) = sqrt(varT / tau_e), fitted as asinh of it
// divided by y_scale (the roots of rare variants are heavy-tailed):
//   t_proposal = sinh(output * y_scale) / sqrt(varT / tau_e).
#pragma once

namespace saige {

constexpr int kMlpIn = 6;
constexpr int kMlpH = 64;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpH;
constexpr int kMlpB2 = kMlpW2 + kMlpH * kMlpH;
constexpr int kMlpW3 = kMlpB2 + kMlpH;
constexpr int kMlpB3 = kMlpW3 + kMlpH;
constexpr int kMlpParams = kMlpB3 + 1;  // 4673

__device__ __forceinline__ float mlp_gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float mlp_gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

}  // namespace saige
