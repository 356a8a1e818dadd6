"""Resident chain-rule Gaussian-boson sampling on the GPU (reference
``sampling/sampler.py:214-244``, ``core/sampler_session.cu``).

The host computes the reduced A-matrices ``A_k = X (I - Q_k^-1)`` of the first
``k`` modes once (``Q_k`` the Husimi matrix of the reduced covariance, reference
``sampling/gbs.py:228-238`` and ``sampling/sampler.py:99-105``); the whole chain
-- prefix state, conditional hafnians (the repeated-row sieve), the inverse-CDF
draw -- then runs on the device and only the samples come back.

The drawn distribution is the sequentially per-mode truncated chain rule: mode
``k`` ~ ``P(n_k | prefix)`` restricted to ``0..cutoff`` and renormalised.
Uniforms come from Philox4x32-10 (counter = (draw, mode), key = seed), so the
samples are reproducible for a seed and distributionally -- not bitwise --
equivalent to any other sampler of the same distribution.
"""
from __future__ import annotations

from math import factorial

import numpy as np

__all__ = ["qmat", "reduced_A", "sample_resident"]


def qmat(cov: np.ndarray, hbar: float = 2.0) -> np.ndarray:
    """Husimi Q covariance (a, a^dagger basis) of an xxpp Wigner covariance."""
    N = len(cov) // 2
    I = np.identity(N)
    x = cov[:N, :N] * 2 / hbar
    xp = cov[:N, N:] * 2 / hbar
    p = cov[N:, N:] * 2 / hbar
    aidaj = (x + p + 1j * (xp - xp.T) - 2 * I) / 4
    aiaj = (x - p + 1j * (xp + xp.T)) / 4
    return np.block([[aidaj, aiaj.conj()], [aiaj, aidaj.conj()]]) + np.identity(2 * N)


def reduced_A(cov: np.ndarray, k: int, hbar: float = 2.0) -> np.ndarray:
    """A-matrix of the state reduced to its first ``k`` modes (2k x 2k)."""
    M = cov.shape[0] // 2
    ix = list(range(k)) + list(range(M, M + k))
    Q = qmat(cov[np.ix_(ix, ix)], hbar)
    X = np.block([[np.zeros((k, k)), np.eye(k)], [np.eye(k), np.zeros((k, k))]])
    return X @ (np.eye(2 * k) - np.linalg.inv(Q))


def sample_resident(cov: np.ndarray, num_samples: int, cutoff: int = 6, hbar: float = 2.0,
                    seed: int | None = None) -> np.ndarray:
    """``num_samples`` photon patterns (int64, shape (num_samples, M)) of the
    zero-displacement Gaussian state ``cov`` (xxpp), drawn entirely on the GPU."""
    from .ext import gbskernels_ext as ext
    cov = np.asarray(cov, dtype=np.float64)
    M = cov.shape[0] // 2
    flat, off, tot = [], [], 0
    for k in range(1, M + 1):
        a = np.ascontiguousarray(reduced_A(cov, k, hbar), dtype=np.complex128).reshape(-1)
        off.append(tot)
        flat.append(a)
        tot += a.size
    ak = np.concatenate(flat) if flat else np.zeros(0, np.complex128)
    invfac = np.array([1.0 / factorial(j) for j in range(cutoff + 1)])
    out = ext.sample_resident(ak, np.array(off, np.int32), M, int(num_samples), int(cutoff), 0, invfac,
                              0 if seed is None else int(seed))
    return out.astype(np.int64)
