"""Shared cohort builders and comparisons for the parity tests."""
import numpy as np


def random_bed(n, m, seed, p_missing=0.0, mafs=None):
    """Random .bed body (m, ceil(n/4)) with HWE genotypes and optional missing codes."""
    rng = np.random.default_rng(seed)
    if mafs is None:
        mafs = rng.uniform(0.02, 0.5, size=m)
    g = rng.binomial(2, mafs[:, None], size=(m, n))
    codes = np.where(g == 2, 0, np.where(g == 1, 2, 3)).astype(np.uint8)
    if p_missing > 0:
        codes[rng.random((m, n)) < p_missing] = 1
    bpm = (n + 3) // 4
    pad = np.zeros((m, bpm * 4), np.uint8)
    pad[:, :n] = codes
    pad = pad.reshape(m, bpm, 4)
    body = (pad[:, :, 0] | (pad[:, :, 1] << 2) | (pad[:, :, 2] << 4) | (pad[:, :, 3] << 6)).astype(np.uint8)
    return body, bpm


def decode_np(body, n):
    """numpy decode of a .bed body: (m, n) f64 with NaN for missing."""
    m, bpm = body.shape
    codes = np.stack([(body >> (2 * t)) & 3 for t in range(4)], axis=2).reshape(m, bpm * 4)[:, :n]
    lut = np.array([2.0, np.nan, 1.0, 0.0])
    return lut[codes]


def neglog10(p):
    return -np.log10(np.clip(p, 1e-300, None))


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
