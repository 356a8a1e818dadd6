"""Pins the oracle's sum|term| indicators (the kappa numerators of the auto tier
and the fp64 error scales of the GPU tests) independently of the oracle:

* closed forms with exact rational arithmetic: on the pair-diagonal matrix
  A = (+)_i [[0, a_i], [a_i, 0]] every power-trace subset term is
  (-1)^(n - |S|) h_n(a_S), h_n the complete homogeneous symmetric polynomial,
  so haf_abs = sum_S h_n(a_S) and haf = prod a_i; on O = a I_2n every
  torontonian subset term is (-1)^(n - |S|) (1 - a)^-|S|, so
  tor_abs = (1 + 1 / (1 - a))^n;
* a second algorithm for every subset term of random complex inputs: the
  generating function det(I - t C)^(-1/2) exp(t/2 d^T X (I - t C)^-1 d) through
  the eigenvalues of C = B_S X and binomial series (the oracle forms power
  traces by repeated products and runs the Newton recurrence on them);
* the torontonian's subset determinants through LAPACK, and multiplicativity
  of tor_abs over direct sums of modes;
* abs >= |value| on random inputs, with equality where every term shares one
  sign.

Reference definitions: cpu_ref/diagnostics.py:74-149 (hafnian_abs_term_sum,
loop_hafnian_abs_term_sum, torontonian_abs_term_sum).
"""
import itertools
from fractions import Fraction
from math import comb

import numpy as np
import pytest

from conftest import rng_complex, rng_symmetric


# ---------------------------------------------------------------- exact forms
def h_complete(vals, n):
    """complete homogeneous symmetric polynomial h_n(vals), exact (Fractions)"""
    tot = Fraction(0)
    for mono in itertools.combinations_with_replacement(range(len(vals)), n):
        p = Fraction(1)
        for i in mono:
            p *= vals[i]
        tot += p
    return tot


def pair_diagonal(a):
    n = len(a)
    A = np.zeros((2 * n, 2 * n))
    for i, x in enumerate(a):
        A[2 * i, 2 * i + 1] = A[2 * i + 1, 2 * i] = x
    return A


PAIR_CASES = [
    [Fraction(1, 2)],
    [Fraction(3, 4), Fraction(5, 4)],
    [Fraction(1, 2), Fraction(3, 2), Fraction(1, 4)],
    [Fraction(7, 8), Fraction(1, 8), Fraction(5, 4), Fraction(3, 2)],
    [Fraction(1, 4), Fraction(3, 4), Fraction(5, 8), Fraction(9, 8), Fraction(2)],
    [Fraction(k, 8) for k in (3, 5, 7, 9, 11, 13)],
]


@pytest.mark.parametrize("a", PAIR_CASES, ids=lambda a: f"n{len(a)}")
def test_haf_abs_pair_diagonal_exact(oracle, a):
    n = len(a)
    A = pair_diagonal([float(x) for x in a])
    want_abs = sum(h_complete([a[i] for i in S], n)
                   for m in range(1, n + 1) for S in itertools.combinations(range(n), m))
    want_val = Fraction(1)
    for x in a:
        want_val *= x
    assert oracle.haf_abs(A)[0] == pytest.approx(float(want_abs), rel=1e-13)
    assert oracle.haf(A)[0] == pytest.approx(float(want_val), rel=1e-12)
    # zero loops: the loop hafnian's terms are the hafnian's
    assert oracle.lhaf_abs(A)[0] == pytest.approx(float(want_abs), rel=1e-13)
    assert oracle.lhaf(A)[0] == pytest.approx(float(want_val), rel=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("a", [0.2, 0.5, -0.75])
def test_tor_abs_aI_closed_form(oracle, n, a):
    O = a * np.eye(2 * n)
    assert oracle.tor_abs(O)[0] == pytest.approx((1.0 + 1.0 / (1.0 - a)) ** n, rel=1e-13)


def test_tor_abs_x_p_asymmetric_closed_form(oracle):
    # O = diag(a on x, b on p): det(I - O_S) = ((1 - a)(1 - b))^|S|
    n, a, b = 4, 0.3, -0.6
    O = np.diag([a] * n + [b] * n)
    r = ((1 - a) * (1 - b)) ** -0.5
    assert oracle.tor_abs(O)[0] == pytest.approx((1 + r) ** n, rel=1e-13)
    assert oracle.tor(O)[0] == pytest.approx((r - 1) ** n, rel=1e-12)


# ------------------------------------------------- second algorithm, random
def series_term(B, n, loop):
    """[t^n] of det(I - t C)^(-1/2) exp(t/2 d^T X (I - t C)^-1 d), C = B X,
    X the pair swap, through the eigen-decomposition of C"""
    s = B.shape[0]
    X = np.zeros((s, s))
    for t in range(s // 2):
        X[2 * t, 2 * t + 1] = X[2 * t + 1, 2 * t] = 1.0
    C = B @ X
    mu, V = np.linalg.eig(C)
    # det(I - tC)^(-1/2) = prod_j sum_k binom(2k, k) / 4^k (mu_j t)^k
    f = np.zeros(n + 1, complex)
    f[0] = 1.0
    binc = np.array([comb(2 * k, k) / 4.0 ** k for k in range(n + 1)])
    for m in mu:
        g = binc * m ** np.arange(n + 1)
        f = np.convolve(f, g)[: n + 1]
    if loop:
        d = np.diag(B).copy()
        w = (d @ X @ V) * np.linalg.solve(V, d)
        # t/2 d^T X (I - tC)^-1 d = sum_{k >= 1} t^k / 2 sum_j w_j mu_j^(k - 1)
        L = np.zeros(n + 1, complex)
        for k in range(1, n + 1):
            L[k] = 0.5 * np.sum(w * mu ** (k - 1))
        E = np.zeros(n + 1, complex)
        E[0] = 1.0
        for k in range(1, n + 1):
            E[k] = sum(i * L[i] * E[k - i] for i in range(1, k + 1)) / k
        f = np.convolve(f, E)[: n + 1]
    return f[n]


def abs_by_series(A, loop):
    A = np.array(A, complex)
    N = A.shape[0]
    if not loop:
        np.fill_diagonal(A, 0.0)
    n = N // 2
    val, tot = 0.0, 0.0
    for m in range(1, n + 1):
        for S in itertools.combinations(range(n), m):
            idx = [p for i in S for p in (2 * i, 2 * i + 1)]
            t = series_term(A[np.ix_(idx, idx)], n, loop)
            val += (-1) ** (n - m) * t
            tot += abs(t)
    return val, tot


@pytest.mark.parametrize("N", [2, 4, 6, 8])
def test_haf_abs_random_complex_by_eigen_series(oracle, N):
    A = rng_symmetric(3, N, 900 + N, 0.5)
    got_abs, got_val = oracle.haf_abs(A), oracle.haf(A)
    for k in range(3):
        val, tot = abs_by_series(A[k], loop=False)
        assert got_abs[k] == pytest.approx(tot, rel=1e-10)
        assert abs(got_val[k] - val) <= 1e-10 * tot


@pytest.mark.parametrize("N", [2, 4, 6, 8])
def test_lhaf_abs_random_complex_by_eigen_series(oracle, N):
    A = rng_symmetric(3, N, 950 + N, 0.5)
    got_abs, got_val = oracle.lhaf_abs(A), oracle.lhaf(A)
    for k in range(3):
        val, tot = abs_by_series(A[k], loop=True)
        assert got_abs[k] == pytest.approx(tot, rel=1e-10)
        assert abs(got_val[k] - val) <= 1e-10 * tot


def test_abs_not_symmetrised_by_loops(oracle):
    # the loop hafnian's abs differs from the hafnian's once loops are present
    A = rng_symmetric(1, 6, 17, 0.5)
    assert abs(oracle.lhaf_abs(A)[0] - oracle.haf_abs(A)[0]) > 1e-3 * oracle.haf_abs(A)[0]


def tor_abs_lapack(O):
    n = O.shape[0] // 2
    tot = 0.0
    for m in range(n + 1):
        for S in itertools.combinations(range(n), m):
            idx = list(S) + [i + n for i in S]
            d = np.linalg.det(np.eye(2 * m) - O[np.ix_(idx, idx)]) if m else 1.0
            tot += abs(1.0 / np.sqrt(complex(d)))
    return tot


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_tor_abs_random_vs_lapack(oracle, n):
    g = np.random.default_rng(300 + n)
    for scale in (0.15, 0.6):
        X = g.uniform(-scale, scale, (2 * n, 2 * n))
        O = (X + X.T) / 2
        assert oracle.tor_abs(O)[0] == pytest.approx(tor_abs_lapack(O), rel=1e-12)
    # complex, non-symmetric O (the complex-LU path of the reference)
    Oc = rng_complex((2 * n, 2 * n), 330 + n, 0.2)
    assert oracle.tor_abs(Oc)[0] == pytest.approx(tor_abs_lapack(Oc), rel=1e-11)


def test_tor_abs_direct_sum_multiplicative(oracle):
    g = np.random.default_rng(5)
    X = g.uniform(-0.3, 0.3, (6, 6))
    X = (X + X.T) / 2          # 3 modes
    Y = g.uniform(-0.3, 0.3, (4, 4))
    Y = (Y + Y.T) / 2          # 2 modes
    # modes (0, 1, 2 | 3, 4) of a 5-mode O in xxpp order
    O = np.zeros((10, 10))
    for blk, modes in ((X, [0, 1, 2]), (Y, [3, 4])):
        k = len(modes)
        pos = modes + [m + 5 for m in modes]
        for i in range(2 * k):
            for j in range(2 * k):
                O[pos[i], pos[j]] = blk[i, j]
    assert oracle.tor_abs(O)[0] == pytest.approx(oracle.tor_abs(X)[0] * oracle.tor_abs(Y)[0], rel=1e-12)


# ----------------------------------------------------------- abs >= |value|
def test_abs_bounds_value_random(oracle):
    H = rng_symmetric(16, 8, 21)
    assert np.all(oracle.haf_abs(H) >= np.abs(oracle.haf(H)) * (1 - 1e-13))
    assert np.all(oracle.lhaf_abs(H) >= np.abs(oracle.lhaf(H)) * (1 - 1e-13))
    g = np.random.default_rng(22)
    X = g.uniform(-0.2, 0.2, (16, 8, 8))
    O = X + X.transpose(0, 2, 1)
    assert np.all(oracle.tor_abs(O) >= np.abs(oracle.tor(O)) * (1 - 1e-13))


def test_abs_equals_value_when_one_term(oracle):
    # N = 2: a single subset; its term is the whole value
    A = rng_symmetric(6, 2, 23)
    assert np.allclose(oracle.haf_abs(A), np.abs(oracle.haf(A)), rtol=1e-15, atol=0)
    assert np.allclose(oracle.lhaf_abs(A), np.abs(oracle.lhaf(A)), rtol=1e-14, atol=0)
    # every term of a pair-diagonal hafnian with n = 1 is a_0 > 0
    assert oracle.haf_abs(pair_diagonal([0.75]))[0] == 0.75
