"""Single-large torontonian (one real O, dim <= 64, split across the GPU):
against the DD / quad subset-determinant oracle at small sizes, closed forms
and block factorisation at the large ones, certified enclosures (fp64 and dd),
the real-domain refusal, and the public API's input contract."""
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


def physical(nm, seed, eta=0.7):
    O = threshold_O(nm, np.random.default_rng(seed), eta=eta)
    return np.ascontiguousarray((O + O.T) / 2)


def direct_sum(O1, O2):
    """xxpp direct sum: modes of O1 first, then O2's."""
    n1, n2 = O1.shape[0] // 2, O2.shape[0] // 2
    n = n1 + n2
    idx1 = np.r_[np.arange(n1), n + np.arange(n1)]
    idx2 = np.r_[n1 + np.arange(n2), n + n1 + np.arange(n2)]
    O = np.zeros((2 * n, 2 * n))
    O[np.ix_(idx1, idx1)] = O1
    O[np.ix_(idx2, idx2)] = O2
    return O


@pytest.mark.parametrize("nm", [1, 2, 3, 5, 7, 9, 11, 12])
def test_matches_oracle_and_certificates_enclose(ext, oracle, nm):
    print(f"tor_single nm={nm}", flush=True)
    O = physical(nm, 700 + nm)
    v = ext.tor_single(O, min(nm, 14))
    exact = oracle.tor_dd(O.astype(np.complex128))[0].real
    scale = oracle.tor_abs(O.astype(np.complex128))[0]
    assert abs(v - exact) <= 64 * nm * 2.0 ** -52 * scale
    vc, bc = ext.tor_single_certified(O, 0)
    assert vc == v and np.isfinite(bc)                     # certified value is the fp64 value
    assert abs(vc - exact) <= bc
    vd, bd = ext.tor_single_ddcertified(O, nm)
    if nm == 1:  # tor = det(I - O)^(-1/2) - 1 cancels below quad resolution: mpmath
        import mpmath
        with mpmath.workdps(60):
            I_O = mpmath.eye(2) - mpmath.matrix(O.tolist())
            q = 1 / mpmath.sqrt(mpmath.det(I_O)) - 1
            assert abs(mpmath.mpf(vd) - q) <= mpmath.mpf(bd)
    else:
        q = oracle.tor_q(O.astype(np.complex128))[0].real if nm <= 9 else exact
        assert abs(vd - q) <= bd + (0.0 if nm <= 9 else 1e-28 * scale)
    # a few binary64 ulps of the value, or the dd resolution of sum |term|
    assert bd <= 4 * 2.0 ** -52 * abs(vd) + 2.0 ** -100 * scale


def test_agrees_with_batched_engine(ext):
    import gbskernels_amd as gk
    for nm in (14, 18, 20):
        O = physical(nm, 800 + nm)
        v, d = gk.tor_single(O, certified=True)
        vb, bb = ext.tor_certified(O[None].astype(np.complex128))
        assert abs(v - vb[0].real) <= d["abs_error_bound"] + bb[0]


@pytest.mark.parametrize("nm", [16, 24, 26, 32])
def test_closed_form_large(ext, nm):
    """tor(a I) = (a / (1 - a))^n. a = 1/2: the value is exactly 1 and the sum
    cancels by 3^n, resolved by the certified-dd tier."""
    print(f"tor_single closed form nm={nm}", flush=True)
    O = 0.5 * np.eye(2 * nm)
    vd, bd = ext.tor_single_ddcertified(O, 14)
    assert abs(vd - 1.0) <= bd and bd <= 1e-12
    if nm <= 26:
        a = 0.125  # small sums: the fp64 certificate is informative here
        O = a * np.eye(2 * nm)
        v, b = ext.tor_single_certified(O, 14)
        exact = (Fraction(a) / (1 - Fraction(a))) ** nm
        assert abs(Fraction(v) - exact) <= Fraction(b)


def test_block_factorisation(ext):
    """tor(O1 (+) O2) = tor(O1) tor(O2), at 24 modes from two 12-mode blocks."""
    O1, O2 = physical(12, 901), physical(12, 902)
    v, b = ext.tor_single_ddcertified(direct_sum(O1, O2), 12)
    v1, b1 = ext.tor_single_ddcertified(O1, 12)
    v2, b2 = ext.tor_single_ddcertified(O2, 12)
    tol = b + abs(v1) * b2 + (abs(v2) + b2) * b1 + 4e-16 * abs(v1 * v2)
    assert abs(v - v1 * v2) <= tol
    assert b <= 1e-12 * abs(v)


def test_off_domain_and_edges(ext):
    import gbskernels_amd as gk
    bad = 1.5 * np.eye(8)
    assert np.isnan(ext.tor_single(bad, 2))
    v, b = ext.tor_single_certified(bad, 2)
    assert np.isnan(v) and np.isinf(b)
    v, b = ext.tor_single_ddcertified(bad, 2)
    assert np.isnan(v) and np.isinf(b)
    with pytest.raises(ValueError):
        gk.tor_single(bad)
    with pytest.raises(ValueError):
        gk.tor_single(bad, certified=True)
    assert ext.tor_single(np.zeros((0, 0)), 0) == 1.0
    with pytest.raises(ValueError):
        ext.tor_single(np.eye(66), 0)
    with pytest.raises(ValueError):
        ext.tor_single(np.eye(6), 4)                        # g > n
    O = physical(6, 5)
    with pytest.raises(ValueError, match="symmetric"):
        gk.tor_single(O + np.triu(np.full(O.shape, 1e-3), 1))
    with pytest.raises(ValueError, match="real-domain"):
        gk.tor_single(O + 1e-3j)
    with pytest.raises(ValueError, match="groups"):
        gk.tor_single(O, groups=9)
    v = gk.tor_single(O)
    vc, d = gk.tor_single(O, certified=True)
    vd, dd = gk.tor_single(O, dd=True)
    assert v == vc and d["tier"] == "certified-fp64" and dd["tier"] == "certified-dd"
    assert abs(v - vd) <= d["abs_error_bound"] + dd["abs_error_bound"]
