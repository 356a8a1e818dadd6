"""Synthetic inputs of the benchmark workloads (numpy only).

The constructions follow the reference's bench inputs in behaviour
(``bench/_inputs.py:184-196,242-249``: Haar unitaries by QR of a complex
Gaussian with fixed phases; lossy Gaussian states) and its threshold-detector
matrix definition (``sampling/gbs.py:241-275``: O = I - Sigma_x^-1 with
Sigma_x = (cov + I) / 2, hbar = 2); the generators themselves are written here.

* :func:`perm_rows` -- permanent inputs: n random rows (with repetition) of the
  first n columns of a Haar-random 64 x 64 unitary (:func:`perm_haar_rows_workload`:
  the permanent benchmark's 65536 x 16x16 + 1024 x 24x24);
* :func:`gbs_adjacency` / :func:`haf_subset` -- complex symmetric
  A = U diag(tanh r) U^T (U Haar 32 x 32, r ~ U[0.5, 1.2]) and its principal
  submatrices; loop hafnians replace the diagonal by gamma ~ CN(0, 0.3^2);
* :func:`threshold_O` -- torontonian inputs of two-mode squeezed pairs through a
  Haar interferometer with loss;
* :func:`served_mix` -- a request stream of mixed function, size and tier with
  Poisson arrivals.
"""

from __future__ import annotations

import numpy as np


def haar_unitary(m: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r).copy()
    ph /= np.abs(ph)
    return q * ph


def threshold_O(nm: int, rng: np.random.Generator, eta: float = 0.7, rmin: float = 0.3,
                rmax: float = 1.0) -> np.ndarray:
    """Real 2nm x 2nm O (xxpp) of two-mode squeezed pairs (modes 2i, 2i+1,
    r ~ U[rmin, rmax]) through a Haar interferometer with loss eta."""
    S = np.eye(2 * nm)
    for i in range(0, nm - 1, 2):
        r = rng.uniform(rmin, rmax)
        ch, sh = np.cosh(r), np.sinh(r)
        a, b = i, i + 1
        T = np.eye(2 * nm)
        T[a, a] = T[b, b] = ch
        T[a, b] = T[b, a] = sh                      # x quadratures
        T[nm + a, nm + a] = T[nm + b, nm + b] = ch
        T[nm + a, nm + b] = T[nm + b, nm + a] = -sh  # p quadratures
        S = T @ S
    U = haar_unitary(nm, rng)
    SU = np.block([[U.real, -U.imag], [U.imag, U.real]])
    S = SU @ S
    cov = eta * (S @ S.T) + (1.0 - eta) * np.eye(2 * nm)
    sig = (cov + np.eye(2 * nm)) / 2.0
    sig = (sig + sig.T) / 2.0
    O = np.eye(2 * nm) - np.linalg.inv(sig)
    return np.ascontiguousarray((O + O.T) / 2.0)


def perm_rows(U: np.ndarray, n: int, rng: np.random.Generator) -> np.ndarray:
    rows = rng.integers(0, U.shape[0], n)
    return np.ascontiguousarray(U[rows][:, :n])


def gbs_adjacency(rng: np.random.Generator, m: int = 32) -> np.ndarray:
    U = haar_unitary(m, rng)
    r = rng.uniform(0.5, 1.2, m)
    A = (U * np.tanh(r)) @ U.T
    # exactly symmetric: the product's rounding leaves A[i, j] and A[j, i] a few
    # ulps apart, which moves the hafnian by ~1e-16 relative depending on which
    # triangle a formula reads
    return (A + A.T) / 2.0


def haf_subset(A: np.ndarray, N: int, rng: np.random.Generator, loop: bool) -> np.ndarray:
    S = np.sort(rng.choice(A.shape[0], N, replace=False))
    M = np.array(A[np.ix_(S, S)])
    if loop:
        M[np.diag_indices(N)] = 0.3 * (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / np.sqrt(2.0)
    return M


def cancellation(func: str, delta: float, rng: np.random.Generator) -> np.ndarray:
    """A matrix whose value is delta times an O(1) factor while its summands stay
    O(1) (a tunable-cancellation family, like the reference's adversarial
    sweep, bench/_inputs.py:60-123):

    * perm: [[1, 2], [-1, 2 + delta]] (+) R, perm = delta perm(R), R 4x4 complex;
    * haf: a 4x4 block whose matchings 2*2 + 1*1 + 1*(delta - 5) sum to delta,
      (+) a symmetric 4x4 R;
    * lhaf: [[1, 3], [3, delta - 3]] (lhaf = delta) (+) a symmetric 4x4 R;
    * tor: one mode, O = a I_2 with a = delta / (1 + delta): tor = a / (1 - a) =
      delta, as the difference det(I - O)^(-1/2) - 1 of O(1) terms.

    Values are computed from the matrix as rounded (exact values come from the
    ref tier, not from these formulas).
    """
    R = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    if func == "perm":
        A = np.zeros((6, 6), complex)
        A[:2, :2] = [[1.0, 2.0], [-1.0, 2.0 + delta]]
        A[2:, 2:] = R
        return A
    if func in ("haf", "lhaf"):
        R = R + R.T
        if func == "haf":
            np.fill_diagonal(R, 0.0)
            Bk = np.array([[0, 2, 1, 1], [2, 0, delta - 5.0, 1], [1, delta - 5.0, 0, 2], [1, 1, 2, 0]], complex)
            A = np.zeros((8, 8), complex)
            A[:4, :4], A[4:, 4:] = Bk, R
            return A
        A = np.zeros((6, 6), complex)
        A[:2, :2] = [[1.0, 3.0], [3.0, delta - 3.0]]
        A[2:, 2:] = R
        return A
    if func == "tor":
        a = delta / (1.0 + delta)
        return np.array([[a, 0.0], [0.0, a]], complex)
    raise ValueError(func)


def geometric_size(rng: np.random.Generator, p: float, lo: int, hi: int) -> int:
    """lo - 1 + Geom(p) on {lo, lo+1, ...}, resampled until <= hi."""
    while True:
        n = lo - 1 + int(rng.geometric(p))
        if n <= hi:
            return n


def served_mix(n_req: int, seed: int, funcs=("perm", "haf", "lhaf"), nmax: int = 24, p: float = 0.15,
               tiers=(("auto", 1.0),), rate: float = 5e4):
    """Request stream: (func, tier, matrix) items and Poisson arrival offsets (s).

    func uniform over ``funcs``; perm size n = 1 + Geom(p) on 2..nmax; haf / lhaf
    size N = 2k with k ~ Geom(p) on 1..nmax/2 (config-3 style submatrices); tor
    mode count nm = k on 1..nmax/2; tier drawn from ``tiers`` (name, weight).
    """
    rng = np.random.default_rng(seed)
    U64 = haar_unitary(64, np.random.default_rng(1))
    A32 = gbs_adjacency(np.random.default_rng(3))
    tor_pool = {nm: [threshold_O(nm, np.random.default_rng(4000 + 16 * nm + j)) for j in range(4)]
                for nm in range(1, nmax // 2 + 1)}
    names = [t for t, _ in tiers]
    w = np.array([x for _, x in tiers], np.float64)
    w /= w.sum()
    items = []
    for _ in range(n_req):
        f = funcs[int(rng.integers(len(funcs)))]
        t = names[int(rng.choice(len(names), p=w))]
        if f == "perm":
            M = perm_rows(U64, geometric_size(rng, p, 2, nmax), rng)
        elif f in ("haf", "lhaf"):
            M = haf_subset(A32, 2 * geometric_size(rng, p, 1, nmax // 2), rng, f == "lhaf")
        else:
            nm = geometric_size(rng, p, 1, nmax // 2)
            M = tor_pool[nm][int(rng.integers(4))]
        items.append((f, t, M))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    return items, arrivals


def perm_haar_rows_workload(take16: int | None = None, take24: int | None = None):
    """Permanent benchmark inputs: 65536 16x16 and 1024 24x24 complex128
    matrices, each n random rows (with repetition) of the first n columns of one
    Haar 64x64 unitary, all drawn from default_rng(1) in that order. take16 /
    take24 keep only the first matrices of each set (same draws)."""
    rng = np.random.default_rng(1)
    U = haar_unitary(64, rng)
    rows16 = rng.integers(0, 64, (65536, 16))
    rows24 = rng.integers(0, 64, (1024, 24))
    m16 = np.ascontiguousarray(U[:, :16][rows16[:take16]])
    m24 = np.ascontiguousarray(U[:, :24][rows24[:take24]])
    return m16, m24


def repeated_ab_base(modes: int = 3, seed: int = 0):
    """Base matrix of the repeated-row sieve benchmark (reference
    bench/repeated_ab.py:45-50): M = 2 modes, A = 0.3 (G + G^T) with G complex
    standard normal from default_rng(seed), loop weights zero."""
    M = 2 * modes
    g = np.random.default_rng(seed)
    A = g.standard_normal((M, M)) + 1j * g.standard_normal((M, M))
    return np.ascontiguousarray(0.3 * (A + A.T)), np.zeros(M, np.complex128)
