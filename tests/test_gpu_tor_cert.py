"""Certified fp64 torontonian on the GPU: the value is bit-identical to the fp64
tier's (same kernels, same split, same reduction order), the bound encloses the
oracle value on every physical input, and inputs the certified path cannot
cover (complex O, non-SPD subsets, near-singular pivots) are refused with an
infinite bound rather than a wrong finite one."""
import sys
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from physics import threshold_O

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


def physical_stack(nm, B, seed, eta=0.7):
    rng = np.random.default_rng(seed)
    return np.stack([threshold_O(nm, rng, eta=eta) for _ in range(B)]).astype(np.complex128)


@pytest.mark.parametrize("nm", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11])
def test_tor_certified_encloses_dd_oracle(ext, oracle, nm):
    print(f"tor certified nm={nm}", flush=True)
    O = physical_stack(nm, 6, 300 + nm)
    v, bound = ext.tor_certified(O)
    assert np.array_equal(v, ext.tor(O))               # value path is the fp64 tier's
    exact = oracle.tor_dd(O)                            # DD subset determinants
    assert np.all(np.isfinite(bound)) and np.all(bound >= 0)
    assert np.all(np.abs(v - exact) <= bound)           # 100% enclosure
    # informative: a few hundred ulps of sum|term| at most
    _, absn = ext.tor_kappa(O)
    assert np.all(bound <= 1e4 * nm * 2.0 ** -52 * absn)


@pytest.mark.parametrize("nm", [12, 16, 20])
def test_tor_certified_large(ext, oracle, nm):
    print(f"tor certified large nm={nm}", flush=True)
    O = physical_stack(nm, 2, 400 + nm)
    v, bound = ext.tor_certified(O)
    assert np.array_equal(v, ext.tor(O))
    ref = oracle.tor_recursive(O.real)
    assert np.all(np.isfinite(bound))
    # the recursive fp64 oracle carries its own rounding error, of the order of
    # the fp64 value's; the enclosure of the exact value is checked against the
    # DD oracle at nm <= 11, here consistency within bound + oracle error
    _, absn = ext.tor_kappa(O)
    assert np.all(np.abs(v.real - ref) <= bound + 64 * nm * 2.0 ** -52 * absn)


def test_tor_certified_closed_form_and_empty(ext):
    for nm in (1, 4, 9):
        a = 0.2
        O = (a * np.eye(2 * nm))[None].astype(np.complex128)
        v, bound = ext.tor_certified(O)
        fa = Fraction(a)                       # tor(a I_2n) = (a / (1 - a))^n exactly
        exact = (fa / (1 - fa)) ** nm
        assert abs(Fraction(v[0].real) - exact) <= Fraction(bound[0]) and v[0].imag == 0
        # sum|term| = (1 + 1 / (1 - a))^n: the bound is a few ulps of it
        assert bound[0] < 16 * nm * 2.0 ** -52 * (1 + 1 / (1 - a)) ** nm
    v, bound = ext.tor_certified(np.zeros((3, 0, 0), np.complex128))
    assert np.all(v == 1) and np.all(bound == 0)


def test_tor_certified_refuses_off_domain(ext):
    g = np.random.default_rng(5)
    # complex O: the value is the fp64 tier's complex fallback; no certificate
    O = physical_stack(3, 2, 77)
    O = O + 0.01j * (lambda X: X + X.T)(g.standard_normal((6, 6)))[None]
    v, bound = ext.tor_certified(O)
    assert np.array_equal(v, ext.tor(O))
    assert np.all(np.isinf(bound))
    # I - O not SPD on some subset: real path breaks down, fallback value, refused
    Obad = (1.5 * np.eye(4))[None].astype(np.complex128)
    v, bound = ext.tor_certified(Obad)
    assert np.array_equal(v, ext.tor(Obad), equal_nan=True)
    assert np.isinf(bound[0])


def test_tor_certified_near_singular_is_refused_or_enclosing(ext, oracle):
    """I - O close to singular: pivots lose relative accuracy; the bound must
    either grow to cover it or refuse."""
    for delta in (1e-6, 1e-10, 1e-14):
        a = 1.0 - delta
        O = (a * np.eye(4))[None].astype(np.complex128)
        v, bound = ext.tor_certified(O)
        fa = Fraction(a)
        exact = (fa / (1 - fa)) ** 2
        if np.isfinite(bound[0]):
            assert abs(Fraction(v[0].real) - exact) <= Fraction(bound[0])


def test_tor_certified_batch_independent(ext):
    O = physical_stack(7, 9, 88)
    v, b = ext.tor_certified(O)
    for i in (0, 4, 8):
        v1, b1 = ext.tor_certified(O[i:i + 1])
        assert v1[0] == v[i] and b1[0] == b[i]


# ---- certified double-double -------------------------------------------------

@pytest.mark.parametrize("nm", [2, 3, 5, 6])
def test_tor_dd_certified_encloses_quad_oracle(ext, oracle, nm):
    print(f"tor dd certified nm={nm}", flush=True)
    O = physical_stack(nm, 4, 500 + nm)
    v, bound = ext.tor_dd_certified(O)
    assert np.array_equal(v, ext.tor_dd(O))            # value path is the dd tier's
    exact = oracle.tor_q(O)                             # __float128 subset determinants
    assert np.all(np.isfinite(bound))
    assert np.all(np.abs(v - exact) <= bound)
    # the bound is dominated by the final rounding to binary64: a few ulps
    assert np.all(bound <= 4 * 2.0 ** -52 * np.abs(v) + 1e-300)


@pytest.mark.parametrize("nm", [9, 12, 16])
def test_tor_dd_certified_large(ext, oracle, nm):
    print(f"tor dd certified large nm={nm}", flush=True)
    O = physical_stack(nm, 2, 600 + nm)
    v, bound = ext.tor_dd_certified(O)
    assert np.array_equal(v, ext.tor_dd(O))
    ref = oracle.tor_dd(O)                              # DD oracle, error << one binary64 ulp
    assert np.all(np.isfinite(bound))
    assert np.all(np.abs(v - ref) <= bound)
    assert np.all(bound <= 4 * 2.0 ** -52 * np.abs(v))


def test_tor_dd_certified_cancellation_and_refusal(ext):
    for delta in (1e-3, 1e-9, 1e-13):
        a = delta                                       # tor(a I_2) = a / (1 - a): cancels
        O = (a * np.eye(2))[None].astype(np.complex128)
        v, bound = ext.tor_dd_certified(O)
        fa = Fraction(a)
        exact = fa / (1 - fa)
        assert abs(Fraction(v[0].real) - exact) <= Fraction(bound[0])
        assert bound[0] <= 2.0 ** -52 * abs(v[0])      # dd resolves the cancellation
    v, bound = ext.tor_dd_certified((1.5 * np.eye(4))[None].astype(np.complex128))
    assert np.isinf(bound[0])                           # I - O not SPD: refused
    v, bound = ext.tor_dd_certified(np.zeros((2, 0, 0), np.complex128))
    assert np.all(v == 1) and np.all(bound == 0)


def test_tor_dd_certified_tiny_value_mpmath():
    """One mode, O = [[0, d], [d, 0]]: tor = 1 / sqrt(1 - d^2) - 1 ~ d^2 / 2,
    below the resolution of __float128 next to 1; checked with mpmath."""
    import mpmath
    for d in (1.2e-17, 3e-9, 0.25):
        O = np.array([[[0.0, d], [d, 0.0]]], np.complex128)
        v, bound = ext_mod().tor_dd_certified(O)
        with mpmath.workdps(80):
            exact = 1 / mpmath.sqrt(1 - mpmath.mpf(d) ** 2) - 1
            assert abs(mpmath.mpf(v[0].real) - exact) <= mpmath.mpf(bound[0])
        # informative: below the dd resolution of the O(1) subset terms
        assert bound[0] <= 2.0 ** -51 * abs(v[0]) + 1e-30


def ext_mod():
    import gbskernels_ext
    return gbskernels_ext
