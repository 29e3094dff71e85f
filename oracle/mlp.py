"""CPU checker for the learned saddle-point root proposal — test infrastructure only.

numpy float64 forward and backward passes of the 6 -> 64 -> 64 -> 1 GELU network with the
parameter layout of csaige_amd/csrc/mlp_common.hpp.  The network has no counterpart in
saige-rs; this restatement is pinned against central finite differences of its own loss
(tests/test_oracle_mlp.py), and the GPU kernels are checked against it.
"""
import numpy as np
from scipy.special import erf

NI, NH = 6, 64
W1 = 0
B1 = W1 + NH * NI
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
NP = B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(NH, NI), p[B1:W2], p[W2:B2].reshape(NH, NH), p[B2:W3], p[W3:B3],
            p[B3])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def standardize(x, norm):
    norm = np.asarray(norm, np.float64)
    return (np.asarray(x, np.float64) - norm[:NI]) * norm[NI:2 * NI]


def forward(p, xn):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    a1 = xn @ w1.T + b1
    h1 = gelu(a1)
    a2 = h1 @ w2.T + b2
    h2 = gelu(a2)
    return h2 @ w3 + b3, (a1, h1, a2, h2)


def target(y, y_scale):
    """The fitted target: asinh of the scaled saddle point over y_scale."""
    return np.arcsinh(np.asarray(y, np.float64)) / y_scale


def loss_grad(p, xn, yn):
    """Mean squared error against yn = target(y) and its gradient (saige_mlp_loss_grad)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    out, (a1, h1, a2, h2) = forward(p, xn)
    n = len(yn)
    err = out - yn
    dout = 2.0 * err / n
    g = np.zeros(NP)
    g[W3:B3] = h2.T @ dout
    g[B3] = dout.sum()
    d2 = dout[:, None] * w3[None, :] * gelu_grad(a2)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ w2) * gelu_grad(a1)
    g[W1:B1] = (d1.T @ xn).ravel()
    g[B1:W2] = d1.sum(0)
    return float(np.mean(err * err)), g
