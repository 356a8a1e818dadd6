"""GBS probability layer on the GPU kernels (reference ``sampling/gbs.py``).

Photon-number probabilities of pure Gaussian states (hafnians), displaced states
(loop hafnians), and threshold-detector click probabilities (marginalised, or
directly from torontonians), with the reference's names, arguments and
conventions. Only ``backend="gpu"`` is served (the reference's CPU loops are its
``cpu_ref``).

MI355X-first evaluation: a pattern's matrix ``B_n`` repeats row/column ``i`` of
the kernel ``n_i`` times, which is exactly the repeated-row loop hafnian's input,
so every pattern's (loop) hafnian comes from ONE sieve launch over the pattern
table (cost ``prod(n_i + 1)`` per pattern, nothing expanded or copied); the
torontonian click probabilities go through one bucketing :class:`Workspace`.
"""
from __future__ import annotations

from itertools import product
from math import cosh, factorial, prod
from typing import Any

import numpy as np

__all__ = ["random_gbs_kernel", "fock_patterns", "probabilities", "total_probability",
           "displaced_probabilities", "displaced_total", "threshold_probabilities", "threshold_total",
           "threshold_O_xxpp", "torontonian_threshold_probabilities"]


def _gpu_only(backend: str) -> None:
    if backend not in ("cpu", "gpu"):
        raise ValueError(f"unknown backend {backend!r}; expected 'cpu' or 'gpu'")
    if backend != "gpu":
        raise NotImplementedError("gbskernels_amd is the GPU backend: the CPU loops are the reference's cpu_ref")


def _fp64_only(precision: str) -> None:
    if precision != "fp64":
        raise NotImplementedError("the sieve-based probability layer is fp64 (the sieve has fp64 and "
                                  "certified tiers; use gbskernels_amd.lhaf_repeated(certified=True))")


def random_gbs_kernel(m: int, seed: int, r_min: float = 0.1, r_max: float = 0.45) -> tuple[np.ndarray, np.ndarray]:
    """Pure-state GBS kernel ``B = U diag(tanh r) U^T`` (Haar ``U``, ``r`` uniform
    in ``[r_min, r_max]``); returns ``(B, r)``."""
    g = np.random.default_rng(seed)
    r = g.uniform(r_min, r_max, m)
    z = (g.standard_normal((m, m)) + 1j * g.standard_normal((m, m))) / np.sqrt(2.0)
    q, rr = np.linalg.qr(z)
    ph = np.diagonal(rr).copy()
    U = q * (ph / np.abs(ph))
    return U @ np.diag(np.tanh(r)) @ U.T, r


def fock_patterns(m: int, cutoff: int) -> list[tuple[int, ...]]:
    """All patterns over ``m`` modes with each occupancy in ``[0, cutoff)``."""
    return list(product(range(cutoff), repeat=m))


def _repeated(A: np.ndarray, gamma, patterns) -> np.ndarray:
    from . import lhaf_repeated
    reps = np.array(patterns, dtype=np.int32).reshape(len(patterns), A.shape[0])
    if len(reps) == 0:
        return np.zeros(0, np.complex128)
    v = lhaf_repeated(A, gamma, reps, backend="gpu")
    if not np.all(np.isfinite(v)):
        raise ValueError("a pattern is beyond the sieve's domain (more than 64 photons or 2^40 terms)")
    return v


def probabilities(B: np.ndarray, r: np.ndarray, cutoff: int, precision: str = "fp64",
                  backend: str = "gpu") -> tuple[list[tuple[int, ...]], np.ndarray]:
    """``P(n) = |haf(B_n)|^2 / (prod n_i! prod cosh r_i)`` for every pattern with
    occupancies below ``cutoff``; odd total photon numbers have probability 0."""
    _gpu_only(backend)
    _fp64_only(precision)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    norm = prod(cosh(float(x)) for x in r)
    patterns = fock_patterns(B.shape[0], cutoff)
    even = [p for p in patterns if sum(p) % 2 == 0]
    hafs = dict(zip(even, _repeated(B, None, even)))
    probs = np.zeros(len(patterns))
    for k, n in enumerate(patterns):
        if sum(n) % 2 == 0:
            probs[k] = abs(hafs[n]) ** 2 / (prod(factorial(x) for x in n) * norm)
    return patterns, probs


def total_probability(B: np.ndarray, r: np.ndarray, cutoff: int, backend: str = "gpu") -> float:
    """Sum of the probabilities up to ``cutoff`` (tends to 1)."""
    return float(probabilities(B, r, cutoff, backend=backend)[1].sum())


def displaced_probabilities(B: np.ndarray, r: np.ndarray, alpha: np.ndarray, cutoff: int,
                            precision: str = "fp64", backend: str = "gpu") -> tuple[list[tuple[int, ...]], np.ndarray]:
    """Displaced pure state: ``P(n) = |pref lhaf(A_n; diag gamma_n)|^2 / (prod n_i! prod cosh r_i)``
    with ``A = -B``, ``gamma = alpha - A conj(alpha)``,
    ``pref = exp(-(|alpha|^2 - alpha^* A alpha^*) / 2)`` (reference
    ``sampling/gbs.py:142-191``); ``alpha = 0`` gives :func:`probabilities`."""
    _gpu_only(backend)
    _fp64_only(precision)
    B = np.asarray(B, dtype=np.complex128)
    alpha = np.asarray(alpha, dtype=np.complex128)
    norm = prod(cosh(float(x)) for x in r)
    A = np.ascontiguousarray(-B)
    gamma = alpha - A @ np.conj(alpha)
    pref = np.exp(-0.5 * (np.vdot(alpha, alpha) - alpha.conj() @ A @ alpha.conj()))
    patterns = fock_patterns(B.shape[0], cutoff)
    lh = _repeated(A, gamma, patterns)
    probs = np.array([float(abs(pref * lh[k]) ** 2 / (prod(factorial(x) for x in n) * norm))
                      for k, n in enumerate(patterns)])
    return patterns, probs


def displaced_total(B: np.ndarray, r: np.ndarray, alpha: np.ndarray, cutoff: int, backend: str = "gpu") -> float:
    return float(displaced_probabilities(B, r, alpha, cutoff, backend=backend)[1].sum())


def threshold_probabilities(B: np.ndarray, r: np.ndarray, cutoff: int, precision: str = "fp64",
                            backend: str = "gpu") -> dict[tuple[int, ...], float]:
    """Click-pattern probabilities by marginalising the photon-number
    distribution up to ``cutoff``."""
    patterns, probs = probabilities(B, r, cutoff, precision, backend=backend)
    out: dict[tuple[int, ...], float] = {}
    for n, p in zip(patterns, probs):
        c = tuple(1 if x > 0 else 0 for x in n)
        out[c] = out.get(c, 0.0) + float(p)
    return out


def threshold_total(B: np.ndarray, r: np.ndarray, cutoff: int, backend: str = "gpu") -> float:
    return float(sum(threshold_probabilities(B, r, cutoff, backend=backend).values()))


def threshold_O_xxpp(cov: np.ndarray, hbar: float = 2.0) -> tuple[np.ndarray, float]:
    """Real torontonian matrix ``Ox = I - Sigma_x^-1`` of a zero-displacement
    state, ``Sigma_x = (cov 2/hbar + I) / 2`` (both symmetrised), and
    ``0.5 log det Sigma_x``; ``P(S) = tor(Ox_S) exp(-log_sqrt_det)`` and ``Ox`` is
    a valid :func:`gbskernels_amd.tor_single` input."""
    n2 = cov.shape[0]
    Sx = (np.asarray(cov, dtype=np.float64) * (2.0 / hbar) + np.eye(n2)) / 2.0
    Sx = np.ascontiguousarray((Sx + Sx.T) / 2.0)
    if float(np.min(np.linalg.eigvalsh(Sx))) <= 0:
        raise ValueError("Husimi covariance is not positive definite")
    sign, logdet = np.linalg.slogdet(Sx)
    Ox = np.eye(n2) - np.linalg.inv(Sx)
    return np.ascontiguousarray((Ox + Ox.T) / 2.0), 0.5 * float(logdet)


def torontonian_threshold_probabilities(cov: np.ndarray, hbar: float = 2.0, precision: str = "fp64",
                                        backend: str = "gpu") -> dict[tuple[int, ...], float]:
    """Click probabilities straight from torontonians: ``P(c) = tor(O_c) / sqrt(det Q)``
    with ``O`` reduced to the clicking modes in both blocks; sums to 1 exactly
    (no photon cutoff). The reference evaluates the complex-basis ``O = I - Q^-1``;
    here the real quadrature-basis ``Ox`` of :func:`threshold_O_xxpp` is used, which
    has the same subset determinants (``O = R Ox R^dagger`` for a fixed block
    unitary ``R`` that commutes with the mode restriction, and ``det Q = det Sigma_x``),
    so every torontonian takes the real-arithmetic kernels. One ragged batch
    through a :class:`Workspace`."""
    _gpu_only(backend)
    from . import Workspace
    m = len(cov) // 2
    Ox, log_sqrt_det = threshold_O_xxpp(cov, hbar)
    patterns = list(product((0, 1), repeat=m))
    mats: list[Any] = []
    for c in patterns:
        S = [i for i in range(m) if c[i]]
        idx = S + [i + m for i in S]
        mats.append(np.ascontiguousarray(Ox[np.ix_(idx, idx)]) if S else np.empty((0, 0)))
    with Workspace(backend="gpu") as ws:
        tors = ws.tor_batched(mats, precision=precision)
    scale = float(np.exp(-log_sqrt_det))
    return {c: float(np.real(t)) * scale for c, t in zip(patterns, tors)}
