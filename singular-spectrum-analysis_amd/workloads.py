"""Synthetic series for the five benchmark shapes.

All noise comes from numpy's Philox4x32 bit generator with a fixed seed, so
every caller (tests, bench, serving replay) draws identical inputs.

  single(N, seed)        x[t] = 0.01 t + 2 sin(2 pi t/50) + sin(2 pi t/12 + 0.3) + N(0, 0.5^2)
  ragged_batch(...)      uneven series: trend + 1-3 sinusoids + noise at SNR ~ U(0, 15) dB
  mssa_channels(...)     16 channels mixing 4 latent sinusoids + quadratic trend + noise
"""
import math

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def single(N, seed=1):
    """The single-series signal family (N=1000 and N=20000 shapes)."""
    t = np.arange(N, dtype=np.float64)
    x = 0.01 * t + 2.0 * np.sin(2 * np.pi * t / 50) + np.sin(2 * np.pi * t / 12 + 0.3)
    return x + _rng(seed).normal(0.0, 0.5, N)


def ragged_series(rng, N):
    t = np.arange(N, dtype=np.float64)
    slope = rng.uniform(-0.02, 0.02)
    clean = slope * t
    for _ in range(int(rng.integers(1, 4))):
        period = rng.uniform(5, 200)
        amp = rng.uniform(0.5, 3)
        phase = rng.uniform(0, 2 * np.pi)
        clean = clean + amp * np.sin(2 * np.pi * t / period + phase)
    snr_db = rng.uniform(0, 15)
    p_sig = float(np.mean((clean - clean.mean()) ** 2)) or 1.0
    sd = math.sqrt(p_sig / (10 ** (snr_db / 10)))
    return clean + rng.normal(0.0, sd, N)


def ragged_batch(count=4096, n_lo=500, n_hi=20000, k_lo=10, k_hi=50, seed=4, dtype=np.float32):
    """Uneven series: N log-uniform in [n_lo, n_hi], L = N // 4, k ~ U{k_lo..k_hi}.
    Returns (list of arrays, Ns, Ls, ks)."""
    rng = _rng(seed)
    Ns = np.floor(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), count))).astype(np.int64)
    Ns = np.clip(Ns, n_lo, n_hi)
    ks = rng.integers(k_lo, k_hi + 1, count)
    xs = [ragged_series(rng, int(n)).astype(dtype) for n in Ns]
    Ls = Ns // 4
    return xs, Ns.astype(np.int32), Ls.astype(np.int32), ks.astype(np.int32)


def mssa_channels(M=16, N=10000, seed=5):
    rng = _rng(seed)
    t = np.arange(N, dtype=np.float64)
    lat = np.stack([np.sin(2 * np.pi * t / p) for p in (7, 30, 90, 365)])
    mix = rng.standard_normal((M, 4))
    trend = 1e-8 * (t - N / 2) ** 2
    return mix @ lat + trend + rng.normal(0.0, 0.3, (M, N))
