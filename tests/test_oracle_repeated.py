"""The oracle's repeated-row sieve (restating cpu_ref/repeated.py:61-122) pinned
without the reference: against brute-force matchings of the expanded matrix
(the reference's own ground truth, tests/test_repeated.py:47-63), closed forms
(all-ones base: telephone numbers and double factorials; diagonal base: a
product of Gaussian moments), and the double/__float128 twins against each
other. Also the public API's argument checks (CPU only, no GPU call)."""
from fractions import Fraction
from math import comb

import numpy as np
import pytest

# tests/test_repeated.py:40-41
PATTERNS = [(0, 0, 0), (1, 1, 0), (2, 1, 1), (2, 0, 2), (3, 2, 1), (1, 1, 1), (4, 3, 2), (5, 0, 1), (2, 2, 2)]


def base(M, seed, cplx=True):
    g = np.random.default_rng(seed)
    A = g.standard_normal((M, M)) + (1j * g.standard_normal((M, M)) if cplx else 0.0)
    A = A + A.T
    gam = g.standard_normal(M) + (1j * g.standard_normal(M) if cplx else 0.0)
    return A, gam


def telephone(N):
    t = [1, 1]
    for k in range(2, N + 1):
        t.append(t[-1] + (k - 1) * t[-2])
    return t[N]


def dfact(k):
    r = 1
    while k > 1:
        r *= k
        k -= 2
    return r


@pytest.mark.parametrize("cplx", [False, True])
def test_sieve_equals_expanded_matchings(oracle, cplx):
    A, gam = base(3, 42, cplx)
    for reps in PATTERNS:
        E = oracle.expand_repeated(A, gam, reps)
        truth = oracle.matchings(E, loop=True)[0] if E.shape[0] else 1.0
        for quad in (False, True):
            got = oracle.lhaf_repeated(A, gam, reps, quad=quad)[0]
            assert abs(got - truth) <= 1e-11 * max(1.0, abs(truth)), (reps, quad)


def test_gamma_zero_is_hafnian_of_expansion(oracle):
    A, _ = base(3, 7)
    for reps in [(2, 2, 0), (2, 1, 1), (4, 2, 2), (3, 1, 0)]:
        E = oracle.expand_repeated(A, None, reps)
        truth = oracle.matchings(E, loop=False)[0]
        got = oracle.lhaf_repeated(A, None, reps)[0]
        assert abs(got - truth) <= 1e-11 * max(1.0, abs(truth)), reps
    # odd total with zero loops: every term carries mu = 0 -> exactly zero
    assert oracle.lhaf_repeated(A, None, (2, 1, 0))[0] == 0.0


def test_all_ones_base_closed_forms(oracle):
    """The expansion of J is J_N: lhaf = telephone number T(N) with unit loops,
    haf = (N-1)!! without."""
    for M, reps in ((2, (3, 4)), (3, (2, 3, 5)), (4, (1, 2, 2, 3)), (5, (2, 2, 2, 2, 2))):
        J = np.ones((M, M))
        N = sum(reps)
        for quad in (False, True):
            v = oracle.lhaf_repeated(J, np.ones(M), reps, quad=quad)[0]
            assert abs(v - telephone(N)) <= 1e-9 * telephone(N), (reps, quad)
            h = oracle.lhaf_repeated(J, None, reps, quad=quad)[0]
            want = dfact(N - 1) if N % 2 == 0 else 0
            assert abs(h - want) <= 1e-9 * max(1, want), (reps, quad)


def gaussian_moment(a, g, n):
    """E[(g + sqrt(a) Z)^n] = sum_k C(n, 2k) (2k-1)!! a^k g^(n-2k) (exact)."""
    return sum(comb(n, 2 * k) * dfact(2 * k - 1) * a ** k * g ** (n - 2 * k) for k in range(n // 2 + 1))


def test_diagonal_base_product_of_moments(oracle):
    """Diagonal A: copies of mode i pair only among themselves (weight a_i) or
    loop (gamma_i), so lhaf = prod_i E[(gamma_i + sqrt(a_i) Z)^(n_i)]."""
    a = [Fraction(3, 4), Fraction(-1, 2), Fraction(5, 3), Fraction(1, 5)]
    g = [Fraction(1, 3), Fraction(2), Fraction(-3, 7), Fraction(1, 2)]
    A = np.diag([float(x) for x in a]).astype(np.complex128)
    gam = np.array([float(x) for x in g], np.complex128)
    for reps in [(1, 0, 0, 0), (2, 1, 0, 3), (4, 4, 1, 2), (6, 0, 5, 3)]:
        exact = Fraction(1)
        for i, n in enumerate(reps):
            exact *= gaussian_moment(a[i], g[i], n)
        for quad in (False, True):
            v = oracle.lhaf_repeated(A, gam, reps, quad=quad)[0]
            assert abs(v.imag) <= 1e-12 * max(1, abs(float(exact)))
            assert abs(v.real - float(exact)) <= 1e-10 * max(1.0, abs(float(exact))), (reps, quad)


def test_double_and_quad_twins_agree(oracle):
    A, gam = base(5, 3)
    reps = np.random.default_rng(8).integers(0, 4, (30, 5)).astype(np.int32)
    d = oracle.lhaf_repeated(A, gam, reps)
    q = oracle.lhaf_repeated(A, gam, reps, quad=True)
    assert np.all(np.abs(d - q) <= 1e-9 * np.maximum(1.0, np.abs(q)))


def test_public_api_argument_checks():
    import gbskernels_amd as gk
    A, gam = base(3, 1)
    assert gk.sieve_term_count((3, 2, 1)) == 24 and gk.sieve_term_count(()) == 1
    with pytest.raises(ValueError, match="reps"):
        gk.lhaf_repeated(A, gam, (1, 2))
    with pytest.raises(ValueError, match="non-negative"):
        gk.lhaf_repeated(A, gam, (1, -1, 0))
    with pytest.raises(ValueError, match="gamma"):
        gk.lhaf_repeated(A, gam[:2], (1, 1, 0))
    with pytest.raises(ValueError, match="square"):
        gk.lhaf_repeated(A[:2], gam, (1, 1, 0))
    with pytest.raises(ValueError, match="finite"):
        gk.lhaf_repeated(np.full((3, 3), np.nan), gam, (1, 1, 0), certified=True)
    with pytest.raises(NotImplementedError):
        gk.lhaf_repeated(A, gam, (1, 1, 0), backend="cpu")


def test_oracle_matches_reference_sieve_checksums(oracle):
    """The reference's published sieve checksums (2048 identical values of the
    pattern (q, ..., q), bench/repeated_ab.py): the quad oracle reproduces each
    value to the reference fp64 kernel's accuracy."""
    import json
    from conftest import GOLDEN
    from gbskernels_amd.workloads import repeated_ab_base
    gold = json.loads((GOLDEN / "repeated_ab_checksums.json").read_text())["rows"]
    A, gam = repeated_ab_base()
    for row in gold:
        ref = complex(*row["checksum"]) / row["batch"]
        q = oracle.lhaf_repeated(A, gam, [row["q"]] * row["M"], quad=True)[0]
        d = oracle.lhaf_repeated(A, gam, [row["q"]] * row["M"])[0]
        assert abs(ref - q) <= 1e-9 * abs(q), row["q"]
        assert abs(d - q) <= 1e-9 * abs(q), row["q"]
