"""Repeated-row loop hafnian (finite-difference sieve) on the GPU against the
oracle's restatement of cpu_ref/repeated.py, the expansion's brute-force
matchings, and size-independent properties at large term counts (closed forms,
block factorisation, mode relabelling, scaling); the certified tier's values are
the plain tier's bit for bit and its bound encloses the exact value."""
import sys
from fractions import Fraction
from math import comb
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu

from test_oracle_repeated import PATTERNS, base, dfact, gaussian_moment, telephone  # noqa: E402


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


@pytest.mark.parametrize("cplx", [False, True])
def test_sieve_matches_expansion_and_oracle(ext, oracle, cplx):
    A, gam = base(3, 42, cplx)
    reps = np.array(PATTERNS, np.int32)
    v = ext.lhaf_repeated(A, gam, reps)
    q = oracle.lhaf_repeated(A, gam, reps, quad=True)
    d = oracle.lhaf_repeated(A, gam, reps)
    for i, r in enumerate(PATTERNS):
        E = oracle.expand_repeated(A, gam, r)
        truth = oracle.matchings(E, loop=True)[0] if E.shape[0] else 1.0
        assert abs(v[i] - truth) <= 1e-11 * max(1.0, abs(truth)), r
        assert abs(v[i] - d[i]) <= 1e-11 * max(1.0, abs(q[i])), r


def test_gamma_zero_and_batched_equals_single(ext, oracle):
    A, _ = base(3, 7)
    z = np.zeros(3, np.complex128)
    reps = np.array([(2, 2, 0), (2, 1, 1), (4, 2, 2), (3, 1, 0), (2, 1, 0)], np.int32)
    v = ext.lhaf_repeated(A, z, reps)
    for i, r in enumerate(reps):
        truth = oracle.matchings(oracle.expand_repeated(A, None, r), loop=False)[0]
        assert abs(v[i] - truth) <= 1e-11 * max(1.0, abs(truth))
        one = ext.lhaf_repeated(A, z, r[None])
        assert one[0] == v[i]                               # batched == looped, bitwise
    assert v[4] == 0.0                                      # odd N, no loops


def test_sampler_shaped_batch_against_oracle(ext, oracle):
    """Many patterns over one base (the chain-rule sampler's shape), M = 8,
    counts 0..4: the plan's inner/outer split varies with the pattern."""
    A, gam = base(8, 11)
    A *= 0.3
    reps = np.random.default_rng(12).integers(0, 4, (120, 8)).astype(np.int32)
    v = ext.lhaf_repeated(A, gam, reps)
    q = oracle.lhaf_repeated(A, gam, reps, quad=True)
    c, bnd = ext.lhaf_repeated_certified(A, gam, reps)
    assert np.array_equal(v.view(np.uint64), c.view(np.uint64))
    assert np.all(np.isfinite(bnd))
    assert np.all(np.abs(v - q) <= bnd)                     # 100% enclosure (quad oracle)
    small = reps.sum(axis=1) <= 12                          # mild cancellation there
    assert np.all(np.abs(v - q)[small] <= 1e-10 * np.maximum(1.0, np.abs(q[small])))


def test_certified_enclosure_reference_case(ext):
    """tests/test_repeated.py:248-282: M = 4 from default_rng(21), the four
    patterns, mpmath loop hafnian of the expansion at 50 digits."""
    import mpmath
    from gbskernels_amd import _ref
    g = np.random.default_rng(21)
    M = 4
    z = (g.standard_normal((M, M)) + 1j * g.standard_normal((M, M))) / np.sqrt(2)
    A = (z + z.T) / 2
    gam = (g.standard_normal(M) + 1j * g.standard_normal(M)) * 0.3
    reps = np.array([[1, 2, 0, 3], [2, 2, 2, 2], [0, 0, 1, 1], [4, 0, 2, 2]], np.int32)
    v, bnd = ext.lhaf_repeated_certified(A, gam, reps)
    assert np.array_equal(v, ext.lhaf_repeated(A, gam, reps))
    for i, row in enumerate(reps):
        idx = np.repeat(np.arange(M), row)
        E = A[np.ix_(idx, idx)].copy()
        np.fill_diagonal(E, gam[idx])
        ex = _ref.lhaf(E, dps=50)
        err = abs(mpmath.mpc(v[i].real, v[i].imag) - mpmath.mpc(ex))
        print(f"pattern {row}: |v| {abs(v[i]):.3e} error {float(err):.2e} bound {bnd[i]:.2e}", flush=True)
        assert err <= bnd[i]
        assert bnd[i] <= 1e-10 * max(1.0, abs(v[i]))


def test_certified_bit_identity_large_N(ext):
    """tests/test_repeated.py:330-349: N = 16, 20, 28."""
    g = np.random.default_rng(44)
    M = 4
    z = (g.standard_normal((M, M)) + 1j * g.standard_normal((M, M))) / np.sqrt(2)
    A = (z + z.T) / 2
    gam = (g.standard_normal(M) + 1j * g.standard_normal(M)) * 0.3
    reps = np.array([[8, 8, 0, 0], [10, 10, 0, 0], [7, 7, 7, 7]], np.int32)
    plain = ext.lhaf_repeated(A, gam, reps)
    cert, bnd = ext.lhaf_repeated_certified(A, gam, reps)
    assert np.array_equal(plain, cert)
    assert np.all(np.isfinite(bnd))


@pytest.mark.parametrize("reps", [(1,) * 20, (2,) * 12, (1, 2) * 7, (3, 3, 3, 3, 3, 3, 2, 2), (5, 4, 3, 2, 1, 1, 1, 1, 2)])
def test_all_ones_closed_form_many_chunks(ext, reps):
    """Term counts 2e4..1e6 (many workgroups, several inner/outer splits): the
    expansion of J is J_N, lhaf = T(N), haf = (N-1)!!. The sieve cancels by
    about e^N here, so N stays near 20."""
    M = len(reps)
    print(f"sieve all-ones reps={reps}", flush=True)
    J = np.ones((M, M), np.complex128)
    r = np.array([reps], np.int32)
    N = sum(reps)
    v, bnd = ext.lhaf_repeated_certified(J, np.ones(M), r)
    exact = telephone(N)
    assert abs(Fraction(v[0].real) - exact) <= Fraction(bnd[0]) and abs(v[0].imag) <= bnd[0]
    assert bnd[0] <= 1e-3 * exact
    h, hb = ext.lhaf_repeated_certified(J, np.zeros(M), r)
    want = dfact(N - 1) if N % 2 == 0 else 0
    assert abs(Fraction(h[0].real) - want) <= Fraction(hb[0])


def test_diagonal_base_product_of_moments(ext):
    a = [Fraction(3, 4), Fraction(-1, 2), Fraction(5, 3), Fraction(1, 5), Fraction(2, 3), Fraction(-1, 3)]
    g = [Fraction(1, 3), Fraction(2), Fraction(-3, 7), Fraction(1, 2), Fraction(1), Fraction(-1, 4)]
    A = np.diag([float(x) for x in a]).astype(np.complex128)
    gam = np.array([float(x) for x in g], np.complex128)
    reps = np.array([(1, 0, 0, 0, 0, 0), (2, 1, 0, 3, 1, 1), (4, 4, 1, 2, 3, 0), (6, 0, 5, 3, 7, 9)], np.int32)
    v, bnd = ext.lhaf_repeated_certified(A, gam, reps)
    for i, r in enumerate(reps):
        exact = Fraction(1)
        for k, n in enumerate(r):
            exact *= gaussian_moment(a[k], g[k], int(n))
        assert abs(Fraction(v[i].real) - exact) <= Fraction(bnd[i]), r
        assert abs(v[i].imag) <= bnd[i]


def test_block_factorisation_relabelling_scaling(ext):
    """Block-diagonal base: lhaf factorises over the blocks; relabelling the
    modes leaves it unchanged; (c^2 A, c gamma) scales it by c^N."""
    A1, g1 = base(4, 5)
    A2, g2 = base(3, 6)
    A1 *= 0.4
    A2 *= 0.4
    A = np.zeros((7, 7), np.complex128)
    A[:4, :4] = A1
    A[4:, 4:] = A2
    gam = np.concatenate([g1, g2])
    n1, n2 = np.array([3, 4, 2, 5], np.int32), np.array([4, 3, 6], np.int32)
    r = np.concatenate([n1, n2])[None]
    v, b = ext.lhaf_repeated_certified(A, gam, r)
    v1, b1 = ext.lhaf_repeated_certified(A1, g1, n1[None])
    v2, b2 = ext.lhaf_repeated_certified(A2, g2, n2[None])
    prod = v1[0] * v2[0]
    tol = b[0] + abs(v1[0]) * b2[0] + (abs(v2[0]) + b2[0]) * b1[0] + 1e-15 * abs(prod)
    assert abs(v[0] - prod) <= tol
    perm = np.array([5, 2, 0, 6, 3, 1, 4])
    vp, bp = ext.lhaf_repeated_certified(A[np.ix_(perm, perm)], gam[perm], r[:, perm])
    assert abs(vp[0] - v[0]) <= b[0] + bp[0]
    c = 0.5  # a power of two: every intermediate scales exactly
    vs = ext.lhaf_repeated(c * c * A, c * gam, r)
    assert vs[0] == c ** int(r.sum()) * v[0]


def test_caps_and_trivial_patterns(ext):
    A, gam = base(3, 1)
    v, b = ext.lhaf_repeated_certified(A, gam, np.array([[0, 0, 0], [40, 25, 0], [1, 0, 0]], np.int32))
    assert v[0] == 1 and b[0] == 0                        # N = 0
    assert np.isnan(v[1].real) and np.isinf(b[1])        # N = 65 > 64
    assert v[2] == gam[0] and np.isfinite(b[2])          # one loop
    # more than 2^40 sieve terms: 3^32 with N = 64
    A32, g32 = base(32, 2)
    v, b = ext.lhaf_repeated_certified(A32, g32, np.full((1, 32), 2, np.int32))
    assert np.isnan(v[0].real) and np.isinf(b[0])
    # M > 32 refuses every pattern; M = 0 is the empty product
    v = ext.lhaf_repeated(np.eye(33), np.zeros(33), np.ones((2, 33), np.int32))
    assert np.all(np.isnan(v.real))
    v = ext.lhaf_repeated(np.zeros((0, 0)), np.zeros(0), np.zeros((3, 0), np.int32))
    assert np.all(v == 1)
    assert ext.lhaf_repeated(A, gam, np.zeros((0, 3), np.int32)).shape == (0,)
    with pytest.raises(ValueError):
        ext.lhaf_repeated(A, gam, np.zeros((2, 4), np.int32))


def test_public_api(ext):
    import gbskernels_amd as gk
    A, gam = base(3, 5)
    reps = np.array([[2, 1, 1], [1, 1, 0], [3, 0, 2]], np.int32)
    vec = gk.lhaf_repeated(A, gam, reps)
    assert vec.shape == (3,)
    for i, r in enumerate(reps):
        assert gk.lhaf_repeated(A, gam, r) == vec[i]
    vc, d = gk.lhaf_repeated(A, gam, reps, certified=True)
    assert np.array_equal(vc, vec) and d["tier"] == "certified-fp64"
    v1, d1 = gk.lhaf_repeated(A, gam, np.array([1, 1, 0]), certified=True)
    assert isinstance(v1, complex) and np.isfinite(d1["abs_error_bound"])
    with pytest.raises(FloatingPointError):
        gk.lhaf_repeated(A, gam, np.array([40, 25, 0]), certified=True)


def test_reference_sieve_checksums(ext, oracle):
    """bench/repeated_ab.py's workload (M = 6, q = 2..6, 2048 identical
    patterns): the batch sum against the reference's published checksum, each
    value enclosed by its certificate around the quad oracle."""
    import json
    from conftest import GOLDEN
    from gbskernels_amd.workloads import repeated_ab_base
    gold = json.loads((GOLDEN / "repeated_ab_checksums.json").read_text())["rows"]
    A, gam = repeated_ab_base()
    for row in gold:
        reps = np.full((row["batch"], row["M"]), row["q"], np.int32)
        v, bnd = ext.lhaf_repeated_certified(A, gam, reps)
        assert np.all(v == v[0]) and np.all(bnd == bnd[0])
        q = oracle.lhaf_repeated(A, gam, reps[0], quad=True)[0]
        ref = complex(*row["checksum"]) / row["batch"]
        assert abs(v[0] - q) <= bnd[0]
        assert abs(v[0] - ref) <= bnd[0] + abs(ref - q)
        chk = complex(np.sum(v))
        assert abs(chk - complex(*row["checksum"])) <= 1e-9 * abs(chk)
