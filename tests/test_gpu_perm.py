"""GPU parity of the Glynn permanent kernels (fp64, kappa, dd) against the oracle,
closed forms and the reference's published checksums. All calls go through the
C ABI (via the gbskernels_ext drop-in module)."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_err, rng_complex

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


EPS = 2.0 ** -52


def glynn_bound(oracle, A):
    """Forward-error scale of an fp64 Glynn evaluation: c * n * eps * sum|term|.
    Random Gaussian matrices cancel (kappa = sum|term| / |perm| ~ 1e2-1e4 at
    n = 16), so agreement is judged against this scale; on well-conditioned
    inputs it is a relative 1e-12 criterion."""
    n = A.shape[-1]
    return 4.0 * max(n, 1) * EPS * oracle.perm_abs(A)


@pytest.mark.parametrize("n", list(range(1, 17)))
def test_perm_fp64_vs_oracle(ext, oracle, n):
    print(f"perm fp64 n={n}", flush=True)
    A = rng_complex((37, n, n), 100 + n)
    got = ext.perm(A)
    exact = oracle.perm_q(A)
    bound = glynn_bound(oracle, A)
    assert np.all(np.abs(got - exact) <= bound)
    # and the oracle's own fp64 value is within the same scale
    assert np.all(np.abs(got - oracle.perm(A)) <= 2 * bound)


@pytest.mark.parametrize("n", [18, 20, 22])
def test_perm_fp64_large_vs_dd(ext, oracle, n):
    print(f"perm fp64 large n={n}", flush=True)
    A = rng_complex((3, n, n), 200 + n, 0.5)
    got = ext.perm(A)
    assert np.all(np.abs(got - oracle.perm_dd(A)) <= glynn_bound(oracle, A))


def haar_rows(n, B, seed):
    g = np.random.default_rng(seed)
    z = (g.standard_normal((64, 64)) + 1j * g.standard_normal((64, 64))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    U = q * (np.diagonal(r) / np.abs(np.diagonal(r)))
    return np.ascontiguousarray(U[g.integers(0, 64, (B, n))][:, :, :n])


def test_perm_well_conditioned_rel_1e12(ext, oracle):
    """Rows of a Haar unitary (the bench workload's family): every instance with
    kappa = sum|term| / |perm| <= 100 is within 1e-12 relative of the quad oracle."""
    for n in (12, 16):
        A = haar_rows(n, 64, 1)
        exact = oracle.perm_q(A)
        kappa = oracle.perm_abs(A) / np.abs(exact)
        ok = kappa <= 100
        assert ok.sum() >= 1, n
        assert rel_err(ext.perm(A[ok]), exact[ok]).max() < 1e-12


@pytest.mark.parametrize("n", [1, 5, 9, 13, 16, 20])
def test_perm_closed_forms(ext, n):
    print(f"perm closed forms n={n}", flush=True)
    J = np.ones((1, n, n))
    assert ext.perm(J)[0] == pytest.approx(math.factorial(n), rel=1e-12)
    assert ext.perm(np.eye(n)[None])[0] == pytest.approx(1.0, abs=1e-12)


def test_perm_known_small_values(ext):
    assert ext.perm(np.array([[[1.0, 2.0], [3.0, 4.0]]]))[0] == pytest.approx(10.0)
    A = np.array([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 10.0]]])
    assert ext.perm(A)[0] == pytest.approx(463.0)


@pytest.mark.parametrize("n", [4, 8, 12, 16])
def test_perm_kappa_value_bitwise_equals_plain(ext, oracle, n):
    print(f"perm kappa n={n}", flush=True)
    A = rng_complex((50, n, n), 300 + n)
    v, a = ext.perm_kappa(A)
    assert np.array_equal(v, ext.perm(A))
    assert rel_err(a, oracle.perm_abs(A)).max() < 1e-10


@pytest.mark.parametrize("n", [3, 8, 12, 16])
def test_perm_dd_vs_oracle(ext, oracle, n):
    print(f"perm dd n={n}", flush=True)
    A = rng_complex((20, n, n), 400 + n)
    assert rel_err(ext.perm_dd(A), oracle.perm_q(A)).max() < 1e-15


def test_perm_dd_survives_cancellation(ext, oracle):
    # rank-deficient-ish: two nearly equal rows -> heavy cancellation in Glynn
    g = np.random.default_rng(9)
    R = g.standard_normal((10, 10)) + 1j * g.standard_normal((10, 10))
    A = R.copy()
    A[1] = A[0] + 1e-9 * R[1]
    A[:, 1] = -A[:, 0] + 1e-9 * R[:, 1]
    want = oracle.perm_q(A)[0]
    dd = ext.perm_dd(A[None])[0]
    fp = ext.perm(A[None])[0]
    assert abs(dd - want) / abs(want) < 1e-14
    assert abs(fp - want) / abs(want) > 1e3 * abs(dd - want) / abs(want)


def test_batch_composition_independence(ext):
    n = 14
    A = rng_complex((64, n, n), 500)
    full = ext.perm(A)
    for idx in ([0], [5, 3], list(range(63, -1, -2))):
        assert np.array_equal(ext.perm(A[idx]), full[idx])
    n = 22  # chunked over workgroups
    A = rng_complex((5, n, n), 501, 0.5)
    full = ext.perm(A)
    assert np.array_equal(ext.perm(A[[3]]), full[[3]])


def test_reference_published_checksums(ext, oracle, ref_checksums):
    for row in ref_checksums:
        if row["func"] not in ("perm", "perm_dd"):
            continue
        dim, B = row["matrix_dim"], row["batch"]
        print(f"checksum {row['func']} dim={dim}", flush=True)
        a = oracle.gen_bench_complex(dim, B, False, 1.0)
        got = float(np.sum(np.abs(getattr(ext, row["func"])(a))))
        assert got == pytest.approx(row["checksum"], rel=1e-12), (row["func"], dim)


def test_session_and_resident(ext, oracle):
    s = ext.Session()
    A = rng_complex((40, 10, 10), 600)
    want = ext.perm(A)
    assert np.array_equal(s.perm(A), want)
    r0 = s.reallocs()
    assert np.array_equal(s.perm(A[:7]), want[:7])
    assert s.reallocs() == r0  # smaller bucket reuses the buffer
    res = s.perm_resident(A)
    assert np.array_equal(res.numpy(), want)
    s.close()


# ------------------------------------------------------------------ certified
@pytest.mark.parametrize("n", [1, 2, 5, 8, 11, 14, 16])
def test_perm_certified_encloses_and_matches_plain(ext, oracle, n):
    print(f"perm certified n={n}", flush=True)
    A = rng_complex((24, n, n), 700 + n)
    v, bound = ext.perm_certified(A)
    assert np.array_equal(v, ext.perm(A))          # value path is the plain kernel's
    exact = oracle.perm_q(A)
    assert np.all(np.isfinite(bound)) and np.all(bound >= 0)
    assert np.all(np.abs(v - exact) <= bound)       # 100% enclosure
    # and the bound is informative: within a modest factor of the forward-error scale
    assert np.all(bound <= 1e3 * glynn_bound(oracle, A) + 1e-300)


@pytest.mark.parametrize("n", [20, 22])
def test_perm_certified_chunked(ext, oracle, n):
    print(f"perm certified chunked n={n}", flush=True)
    A = rng_complex((2, n, n), 750 + n, 0.5)
    v, bound = ext.perm_certified(A)
    assert np.array_equal(v, ext.perm(A))
    assert np.all(np.abs(v - oracle.perm_dd(A)) <= bound)


def test_perm_certified_adversarial_cancellation(ext, oracle):
    """A (+) R with A = [[1, 1], [1, -1 + d]]: the Gray walk drives one row sum
    between d and 2 - d, so the fp64 value loses ~u / d relative accuracy even
    though sum|term| / |perm| is small. The certificate must see it."""
    g = np.random.default_rng(12)
    R = g.standard_normal((7, 7)) + 1j * g.standard_normal((7, 7))
    A = np.zeros((9, 9), complex)
    A[:2, :2] = [[1, 1], [1, -1 + 1e-8]]
    A[2:, 2:] = R
    v, bound = ext.perm_certified(A[None])
    exact = oracle.perm_q(A)[0]
    assert abs(v[0] - exact) <= bound[0]
    assert abs(v[0] - exact) / abs(exact) > 1e-10      # fp64 really is off here
    assert bound[0] / abs(exact) > 1e-9                 # and the bound says so
    assert abs(ext.perm_dd(A[None])[0] - exact) / abs(exact) < 1e-15


def test_perm_certified_refuses_overflow(ext):
    A = np.full((1, 12, 12), 1e30 + 0j)
    v, bound = ext.perm_certified(A)
    assert not np.isfinite(bound[0]) or not np.isfinite(v[0])


# ------------------------------------------------------------ certified dd
@pytest.mark.parametrize("n", [1, 2, 5, 8, 11, 14])
def test_perm_dd_certified_encloses_quad_oracle(ext, oracle, n):
    print(f"perm dd certified n={n}", flush=True)
    A = rng_complex((16, n, n), 900 + n)
    v, bound = ext.perm_dd_certified(A)
    assert np.array_equal(v, ext.perm_dd(A))       # value path is the dd tier's
    exact = oracle.perm_q(A)
    assert np.all(np.isfinite(bound))
    assert np.all(np.abs(v - exact) <= bound)       # 100% enclosure
    # dominated by the final rounding to binary64 (1-norm): a few ulps
    assert np.all(bound <= 4 * EPS * (np.abs(v.real) + np.abs(v.imag)) + 1e-300)


@pytest.mark.parametrize("n", [16, 20])
def test_perm_dd_certified_chunked(ext, oracle, n):
    print(f"perm dd certified chunked n={n}", flush=True)
    A = rng_complex((2, n, n), 950 + n, 0.5)
    v, bound = ext.perm_dd_certified(A)
    assert np.array_equal(v, ext.perm_dd(A))
    assert np.all(np.abs(v - oracle.perm_dd(A)) <= bound)
    assert np.all(bound <= 4 * EPS * (np.abs(v.real) + np.abs(v.imag)))


def test_perm_dd_certified_resolves_cancellation(ext, oracle):
    """The adversarial walk of the fp64 test: dd certifies a tight bound."""
    g = np.random.default_rng(12)
    R = g.standard_normal((7, 7)) + 1j * g.standard_normal((7, 7))
    A = np.zeros((9, 9), complex)
    A[:2, :2] = [[1, 1], [1, -1 + 1e-8]]
    A[2:, 2:] = R
    v, bound = ext.perm_dd_certified(A[None])
    exact = oracle.perm_q(A)[0]
    assert abs(v[0] - exact) <= bound[0]
    assert bound[0] / abs(exact) < 1e-14
    v0, b0 = ext.perm_dd_certified(np.zeros((3, 0, 0), complex))
    assert np.all(v0 == 1) and np.all(b0 == 0)


# ------------------------------------------------------- tuned chunk shape
@pytest.mark.parametrize("family,n,k", [(0, 20, 6), (0, 20, 13), (0, 24, 12), (1, 16, 3), (1, 18, 10)])
def test_perm_split_entry_changes_only_rounding(ext, oracle, family, n, k):
    """Any entry of the chunk-shape table gives the permanent to rounding, with
    the family's tiers still bit-identical (fp64 = kappa value = certified value;
    dd = certified-dd value) and the certificates enclosing the oracle."""
    import ctypes
    from gbskernels_amd import _lib
    L = _lib.lib()
    print(f"perm split family={family} n={n} k={k}", flush=True)
    A = rng_complex((3, n, n), 1200 + n, 0.5)
    ref = oracle.perm_dd(A)
    old = ctypes.c_int()
    assert L.gbs_perm_get_lane_log2(family, n, ctypes.byref(old)) == 0
    base = ext.perm(A) if family == 0 else ext.perm_dd(A)
    try:
        assert L.gbs_perm_set_lane_log2(family, n, k) == 0
        if family == 0:
            v = ext.perm(A)
            assert np.max(rel_err(v, ref)) <= 1e-11 and np.max(rel_err(base, ref)) <= 1e-11
            kv, _ = ext.perm_kappa(A)
            cv, cb = ext.perm_certified(A)
            assert np.array_equal(kv, v) and np.array_equal(cv, v)
            assert np.all(np.abs(cv - ref) <= cb)
        else:
            v = ext.perm_dd(A)
            assert np.max(rel_err(v, ref)) <= 4 * EPS
            cv, cb = ext.perm_dd_certified(A)
            assert np.array_equal(cv, v)
            assert np.all(np.abs(cv - oracle.perm_q(A)) <= cb + 0.0)
    finally:
        assert L.gbs_perm_set_lane_log2(family, n, old.value) == 0
    # back on the built-in entry: the built-in values again, bit for bit
    assert np.array_equal(ext.perm(A) if family == 0 else ext.perm_dd(A), base)


@pytest.mark.parametrize("n", [14, 24, 26, 28])
def test_perm_largest_sizes_closed_forms(ext, n):
    """The sizes whose chunk shape the tuned table moved (fp64 n = 28, dd n = 14
    and 24..28) and their neighbours, against closed forms: a generalised
    permutation matrix P D has perm = prod(d) (every Glynn term has that
    magnitude), and perm(J_n) = n! in double-double (terms up to n^n, exact
    value n! to dd resolution)."""
    print(f"perm largest sizes n={n}", flush=True)
    g = np.random.default_rng(3000 + n)
    d = np.exp(1j * g.uniform(0, 2 * np.pi, n)) * g.uniform(0.8, 1.25, n)
    PD = np.zeros((n, n), complex)
    PD[np.arange(n), g.permutation(n)] = d
    want = np.prod(d)
    for v in (ext.perm(PD[None])[0], ext.perm_dd(PD[None])[0]):
        assert abs(v - want) <= 1e-12 * abs(want)
    v, b = ext.perm_certified(PD[None])
    assert abs(v[0] - want) <= b[0] + 1e-14 * abs(want)  # want: numpy rounding
    f = float(math.factorial(n))
    assert abs(ext.perm_dd(np.ones((1, n, n)))[0] - f) <= 4 * EPS * f
