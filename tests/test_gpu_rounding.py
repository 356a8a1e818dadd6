"""The certified fp64 tier's arithmetic model on the device: every division and
square root it charges one rounding of u|result| must be correctly rounded
(IEEE-754 binary64, round to nearest even). gbs_check_div_sqrt evaluates the
very expressions the certified torontonian (RCERT::div, RCERT::sqrt_) and the
certified hafnian's Newton step (division by the index) use; numpy's float64
division and square root are correctly rounded on x86-64 (SSE2 divsd/sqrtsd,
no flush-to-zero), so the device results must match them bit for bit.

The torontonian's trailing walk takes its pivots' reciprocal square roots from
the hardware estimate refined by one Newton step (gbs_check_rsqrt); the
certified tier charges that 2u|result| on [2^-1000, 2^1000] (and refuses
pivots outside it), checked here against 1 / sqrt(a) in x87 extended precision
(64-bit significand)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TINY = np.finfo(np.float64).tiny          # smallest normal
SUB = np.float64(5e-324)                  # smallest subnormal


def _device(a, b):
    from gbskernels_amd import _lib
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = a.size
    q, s, h = (np.empty(n, np.float64) for _ in range(3))
    _lib.check(_lib.lib().gbs_check_div_sqrt(_lib.ptr(a), _lib.ptr(b), ctypes.c_longlong(n),
                                             _lib.ptr(q), _lib.ptr(s), _lib.ptr(h)), "gbs_check_div_sqrt")
    return q, s, h


def _random_operands(rng, n):
    """Significands uniform in [1, 2), exponents over the whole normal range and
    into the subnormals, random signs."""
    m = 1.0 + rng.random(n)
    e = rng.integers(-1074, 1000, n)
    x = np.ldexp(m, e)                     # ldexp rounds into the subnormal range
    x[x == 0] = SUB
    return np.where(rng.random(n) < 0.5, -x, x)


def _hard_cases(rng, n):
    """Quotients and roots within an ulp of a rounding midpoint: a = m * b (and
    a = m^2) for m halfway between two neighbouring doubles, rounded through
    long double, so the exact a / b (sqrt a) sits next to the midpoint."""
    q = 1.0 + rng.random(n)
    mid = (q.astype(np.longdouble) + np.spacing(q).astype(np.longdouble) / 2)
    b = 1.0 + rng.random(n)
    a_div = (mid * b.astype(np.longdouble)).astype(np.float64)
    scale = np.ldexp(1.0, rng.integers(-500, 500, n))
    a_sqrt = (mid * mid).astype(np.float64)
    return a_div * scale, b * scale, a_sqrt


def _edge_cases():
    pw = np.ldexp(1.0, np.arange(-1074, 1024, 7))
    vals = np.concatenate([pw, pw * 3, [1.0, 2.0, 3.0, 7.0, 10.0, 0.1, TINY, TINY * 1.5, SUB, 3 * SUB, 2 ** 52 * SUB,
                                       np.finfo(np.float64).max, 1 - 2 ** -53, 1 + 2 ** -52]])
    a = np.repeat(vals, vals.size)
    b = np.tile(vals, vals.size)
    return a, b


def _check(a, b):
    q, s, h = _device(a, b)
    with np.errstate(all="ignore"):
        want_q = a / b
        want_s = np.sqrt(np.abs(a))
        want_h = a / (1.0 + np.arange(a.size) % 32)
    for got, want, what in ((q, want_q, "a / b"), (s, want_s, "sqrt|a|"), (h, want_h, "a / j")):
        bad = np.nonzero(got.view(np.int64) != want.view(np.int64))[0]
        assert bad.size == 0, (f"{what}: {bad.size} results differ from the correctly rounded one, e.g. "
                               f"a={a[bad[0]]!r} b={b[bad[0]]!r} got={got[bad[0]]!r} want={want[bad[0]]!r}")


def test_random_operands_correctly_rounded():
    rng = np.random.default_rng(20261203)
    n = 10_000_000
    _check(_random_operands(rng, n), _random_operands(rng, n))


def test_near_midpoint_operands_correctly_rounded():
    rng = np.random.default_rng(7)
    a_div, b, a_sqrt = _hard_cases(rng, 2_000_000)
    _check(a_div, b)
    _check(a_sqrt, np.ones_like(a_sqrt))


def test_edge_operands_correctly_rounded():
    a, b = _edge_cases()
    _check(a, b)
    _check(-a, b)


U = 2.0 ** -53


def _rsqrt(a):
    from gbskernels_amd import _lib
    a = np.ascontiguousarray(a, np.float64)
    r = np.empty(a.size, np.float64)
    _lib.check(_lib.lib().gbs_check_rsqrt(_lib.ptr(a), ctypes.c_longlong(a.size), _lib.ptr(r)), "gbs_check_rsqrt")
    return r


def _check_rsqrt(a):
    """worst relative error, in u, over the operands in [2^-1000, 2^1000]"""
    assert np.finfo(np.longdouble).nmant >= 63, "needs x87 extended precision"
    a = a[(a >= 2.0 ** -1000) & (a <= 2.0 ** 1000)]
    r = _rsqrt(a)
    exact = 1.0 / np.sqrt(a.astype(np.longdouble))
    rel = np.abs((r.astype(np.longdouble) - exact) / exact)
    worst = float(rel.max())
    assert worst <= 2 * U, f"1/sqrt: relative error {worst / U:.3f} u > 2u at a={a[np.argmax(rel)]!r}"
    return worst / U


def test_rsqrt_within_two_ulp():
    rng = np.random.default_rng(53)
    n = 10_000_000
    # pivots of I - O_S live in (0, 2]: dense there, then the whole range
    piv = rng.uniform(0.0, 2.0, n // 2)
    wide = np.ldexp(1.0 + rng.random(n // 2), rng.integers(-1074, 1024, n // 2))
    worst = _check_rsqrt(np.concatenate([piv, wide]))
    assert worst > 0.3  # the check sees the refinement's rounding (not a copy of the reference)


def test_rsqrt_edges():
    pw = np.ldexp(1.0, np.arange(-1074, 1024))
    with np.errstate(over="ignore"):
        near = np.concatenate([pw * (1 + 2.0 ** -52), pw * (1 - 2.0 ** -53), pw * 1.5, pw * 3])
    ends = np.array([2.0 ** -1000, 2.0 ** 1000, np.nextafter(2.0 ** -1000, 1.0), np.nextafter(2.0 ** 1000, 0.0)])
    _check_rsqrt(np.concatenate([pw, near[np.isfinite(near)], ends]))
    # non-positive and NaN operands give a non-finite result (such pivots are
    # flagged before their reciprocal root matters; zero gives NaN: 0 * inf)
    r = _rsqrt(np.array([-1.0, -2.0 ** -600, np.nan, 0.0]))
    assert not np.isfinite(r).any()
