"""Fail-closed certified kernels on non-finite input, called straight through
the C ABI (the public API rejects such input before any launch; these calls
bypass that check). An instance holding a NaN or an infinity has no exact value
to enclose, so its certificate must be refused (non-finite bound); the other
instances of the same batch keep the bits they have when evaluated alone."""
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import rng_complex, rng_symmetric
from physics import threshold_O

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


BAD = [np.nan, np.inf, -np.inf]


def _check(fn, stack, bad_idx):
    v, b = fn(stack)
    assert not np.any(np.isfinite(b[bad_idx])), (b[bad_idx], v[bad_idx])
    clean = [i for i in range(stack.shape[0]) if i not in bad_idx]
    v0, b0 = fn(stack[clean])
    assert np.all(np.isfinite(b0))
    assert np.array_equal(v[clean], v0, equal_nan=True) and np.array_equal(b[clean], b0)


@pytest.mark.parametrize("tier", ["certified", "dd_certified"])
@pytest.mark.parametrize("n", [3, 8, 13])
@pytest.mark.parametrize("bad", BAD)
def test_perm_certificate_refuses_nonfinite_entry(ext, tier, n, bad):
    print(f"perm {tier} n={n} entry={bad}", flush=True)
    A = rng_complex((4, n, n), 900 + n, 0.5)
    A[1, n // 2, n - 1] = bad
    A[3, 0, 0] = complex(0.0, bad)
    _check(getattr(ext, f"perm_{tier}"), A, [1, 3])


@pytest.mark.parametrize("func", ["haf", "lhaf"])
@pytest.mark.parametrize("tier", ["certified", "dd_certified"])
@pytest.mark.parametrize("N", [4, 8, 14])
@pytest.mark.parametrize("bad", BAD)
def test_hafnian_certificate_refuses_nonfinite_entry(ext, func, tier, N, bad):
    print(f"{func} {tier} N={N} entry={bad}", flush=True)
    A = rng_symmetric(4, N, 950 + N, 0.5)
    A[1, 0, N - 1] = A[1, N - 1, 0] = bad
    A[2, N // 2, N // 2] = bad  # a diagonal entry: the loop hafnian reads it
    bad_idx = [1, 2] if func == "lhaf" else [1]
    if func == "haf":
        # the ordinary hafnian zeroes the diagonal, so instance 2 stays valid
        v, b = getattr(ext, f"{func}_{tier}")(A)
        assert not np.isfinite(b[1])
        assert np.isfinite(b[2]) and np.isfinite(v[2])
        A2 = A.copy()
        A2[2, N // 2, N // 2] = 0.0
        v2, b2 = getattr(ext, f"{func}_{tier}")(A2[2:3])
        assert v[2] == v2[0] and b[2] == b2[0]
    _check(getattr(ext, f"{func}_{tier}"), A, bad_idx)


@pytest.mark.parametrize("tier", ["certified", "dd_certified"])
@pytest.mark.parametrize("nm", [3, 6])
@pytest.mark.parametrize("bad", BAD)
def test_tor_certificate_refuses_nonfinite_entry(ext, tier, nm, bad):
    print(f"tor {tier} nm={nm} entry={bad}", flush=True)
    rng = np.random.default_rng(970 + nm)
    O = np.stack([threshold_O(nm, rng) for _ in range(4)]).astype(np.complex128)
    d = 2 * nm
    O[1, 0, d - 1] = O[1, d - 1, 0] = bad
    O[3, d // 2, d // 2] = bad
    _check(getattr(ext, f"tor_{tier}"), O, [1, 3])


@pytest.mark.parametrize("bad", BAD)
def test_sieve_certificate_refuses_nonfinite_input(ext, bad):
    print(f"sieve certified entry={bad}", flush=True)
    M = 5
    A = rng_symmetric(1, M, 990, 0.5)[0]
    gamma = rng_complex((M,), 991, 0.3)
    reps = np.array([[1, 0, 2, 1, 0], [0, 1, 1, 0, 2], [2, 1, 0, 0, 1]], np.int32)
    v0, b0 = ext.lhaf_repeated_certified(A, gamma, reps)
    assert np.all(np.isfinite(b0))
    Ab = A.copy()
    Ab[0, 2] = Ab[2, 0] = bad  # modes 0 and 2 occur together only in the first pattern
    v, b = ext.lhaf_repeated_certified(Ab, gamma, reps)
    assert not np.isfinite(b[0])
    # a pattern that never reads the bad entry is either unchanged or refused,
    # never certified with other bits
    for k in (1, 2):
        assert not np.isfinite(b[k]) or (v[k] == v0[k] and b[k] == b0[k])
    gb = gamma.copy()
    gb[4] = bad  # mode 4 occurs in the second and third patterns
    v, b = ext.lhaf_repeated_certified(A, gb, reps)
    assert not np.isfinite(b[1]) and not np.isfinite(b[2])
    assert not np.isfinite(b[0]) or (v[0] == v0[0] and b[0] == b0[0])
