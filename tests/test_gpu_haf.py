"""GPU parity of the power-trace hafnian / loop hafnian kernels (fp64, kappa,
dd) against the oracle (quad and DD tiers), closed forms and the reference's
published checksums, through the C ABI (gbskernels_ext drop-in)."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_err, rng_symmetric

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu
EPS = 2.0 ** -52


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


def pt_bound(abs_sum, N):
    """Error scale of an fp64 power-trace evaluation: 256 N eps sum_S |term_S|."""
    return 256.0 * max(N, 1) * EPS * abs_sum


@pytest.mark.parametrize("N", [2, 4, 6, 8, 10, 12, 14, 16])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_fp64_vs_quad(ext, oracle, func, N):
    print(f"{func} fp64 N={N}", flush=True)
    A = rng_symmetric(9 if N <= 12 else 3, N, 10 * N + (func == "lhaf"))
    got = getattr(ext, func)(A)
    exact = getattr(oracle, f"{func}_q")(A)
    bound = pt_bound(getattr(oracle, f"{func}_abs")(A), N)
    assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize("N", [18, 20, 24])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_fp64_large_vs_dd_oracle(ext, oracle, func, N):
    print(f"{func} fp64 large N={N}", flush=True)
    A = rng_symmetric(2, N, 30 * N + (func == "lhaf"), 0.5)
    got = getattr(ext, func)(A)
    exact = getattr(oracle, f"{func}_dd")(A)
    bound = pt_bound(getattr(oracle, f"{func}_abs")(A), N)
    assert np.all(np.abs(got - exact) <= bound)


def test_config1_hafnian_well_conditioned(ext, oracle):
    """4096 real symmetric 8x8 A = G + G^T (default_rng(0)): relative error
    <= 1e-12 against the quad oracle wherever kappa <= 100."""
    G = np.random.default_rng(0).standard_normal((4096, 8, 8))
    A = G + G.transpose(0, 2, 1)
    got = ext.haf(A)
    exact = oracle.haf_q(A)
    kappa = oracle.haf_abs(A) / np.abs(exact)
    ok = kappa <= 100
    assert ok.mean() > 0.5
    assert rel_err(got[ok], exact[ok]).max() < 1e-12
    assert np.all(np.abs(got - exact) <= pt_bound(oracle.haf_abs(A), 8))


@pytest.mark.parametrize("n", range(1, 10))
def test_haf_all_ones(ext, n):
    J = np.ones((1, 2 * n, 2 * n))
    want = math.prod(range(2 * n - 1, 0, -2))
    assert ext.haf(J)[0] == pytest.approx(want, rel=1e-12)
    assert ext.haf_dd(J)[0] == pytest.approx(want, rel=1e-15)


def test_lhaf_closed_forms(ext):
    a, b, c = 2.0, 3.0, 5.0
    assert ext.lhaf(np.array([[[a, b], [b, c]]]))[0] == pytest.approx(b + a * c)
    d = np.array([2.0, 3.0, 5.0, 7.0, 1.5, -2.0])
    assert ext.lhaf(np.diag(d)[None])[0] == pytest.approx(np.prod(d))
    telephone = [1, 1, 2, 4, 10, 26, 76, 232, 764, 2620, 9496, 35696, 140152, 568504, 2390480, 10349536,
                 46206736]
    for N in range(2, 17, 2):
        assert ext.lhaf(np.ones((1, N, N)))[0] == pytest.approx(telephone[N], rel=1e-12)
        assert ext.lhaf_dd(np.ones((1, N, N)))[0] == pytest.approx(telephone[N], rel=1e-15)


def test_trivial_sizes(ext):
    assert ext.haf(np.zeros((2, 0, 0)))[0] == 1.0
    assert ext.lhaf(np.zeros((2, 0, 0)))[0] == 1.0
    assert ext.haf(np.ones((2, 3, 3)))[0] == 0.0
    assert np.isnan(ext.lhaf(np.ones((1, 3, 3)))[0])  # odd N is augmented by the API layer


def test_haf_ignores_diagonal_lhaf_uses_it(ext, oracle):
    A = rng_symmetric(4, 10, 77)
    D = A.copy()
    for b in range(4):
        np.fill_diagonal(D[b], 50.0)
    assert rel_err(ext.haf(D), ext.haf(A)).max() < 1e-12
    assert rel_err(ext.lhaf(D), oracle.lhaf_q(D)).max() < 1e-11


@pytest.mark.parametrize("N", [4, 8, 12])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_kappa_value_bitwise_and_abs(ext, oracle, func, N):
    A = rng_symmetric(16, N, 5 * N)
    v, a = getattr(ext, f"{func}_kappa")(A)
    assert np.array_equal(v, getattr(ext, func)(A))
    assert rel_err(a, getattr(oracle, f"{func}_abs")(A)).max() < 1e-9


@pytest.mark.parametrize("N", [2, 6, 10, 12])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_dd_vs_quad(ext, oracle, func, N):
    print(f"{func} dd N={N}", flush=True)
    A = rng_symmetric(6, N, 3 * N + 1)
    got = getattr(ext, f"{func}_dd")(A)
    exact = getattr(oracle, f"{func}_q")(A)
    assert np.all(np.abs(got - exact) <= 1e-14 * np.abs(exact) + 1e-28 * getattr(oracle, f"{func}_abs")(A))


def test_batch_composition_independence(ext):
    for func, N in (("haf", 12), ("lhaf", 10)):
        A = rng_symmetric(40, N, 900 + N)
        full = getattr(ext, func)(A)
        for idx in ([0], [7, 2], list(range(39, -1, -3))):
            assert np.array_equal(getattr(ext, func)(A[idx]), full[idx])


def test_reference_published_checksums(ext, oracle, ref_checksums):
    for row in ref_checksums:
        if row["func"] not in ("haf", "haf_dd", "lhaf", "lhaf_dd"):
            continue
        dim, B = row["matrix_dim"], row["batch"]
        print(f"checksum {row['func']} dim={dim}", flush=True)
        a = oracle.gen_bench_complex(dim, B, True, 1.0)
        got = float(np.sum(np.abs(getattr(ext, row["func"])(a))))
        assert got == pytest.approx(row["checksum"], rel=1e-11), (row["func"], dim)


# ------------------------------------------------------------------ certified
@pytest.mark.parametrize("N", [2, 4, 8, 12, 16])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_certified_encloses_and_matches_plain(ext, oracle, func, N):
    print(f"{func} certified N={N}", flush=True)
    A = rng_symmetric(8 if N <= 12 else 3, N, 40 * N + (func == "lhaf"))
    v, bound = getattr(ext, f"{func}_certified")(A)
    assert np.array_equal(v, getattr(ext, func)(A))
    exact = getattr(oracle, f"{func}_q")(A)
    assert np.all(np.isfinite(bound)) and np.all(bound >= 0)
    assert np.all(np.abs(v - exact) <= bound)
    # informative: within a modest factor of the power-trace error scale
    assert np.all(bound <= 1e4 * pt_bound(getattr(oracle, f"{func}_abs")(A), N))


@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_certified_large_encloses_dd_oracle(ext, oracle, func):
    A = rng_symmetric(2, 20, 55 + (func == "lhaf"), 0.5)
    v, bound = getattr(ext, f"{func}_certified")(A)
    assert np.array_equal(v, getattr(ext, func)(A))
    assert np.all(np.abs(v - getattr(oracle, f"{func}_dd")(A)) <= bound)


def test_certified_config1_relative_bound(ext, oracle):
    G = np.random.default_rng(0).standard_normal((512, 8, 8))
    A = G + G.transpose(0, 2, 1)
    v, bound = ext.haf_certified(A)
    exact = oracle.haf_q(A)
    assert np.all(np.abs(v - exact) <= bound)
    assert np.median(bound / np.abs(exact)) < 1e-11


# ------------------------------------------------------------ certified dd
@pytest.mark.parametrize("N", [2, 4, 8, 12])
@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_dd_certified_encloses_quad_oracle(ext, oracle, func, N):
    print(f"{func} dd certified N={N}", flush=True)
    A = rng_symmetric(6, N, 70 * N + (func == "lhaf"))
    v, bound = getattr(ext, f"{func}_dd_certified")(A)
    assert np.array_equal(v, getattr(ext, f"{func}_dd")(A))   # the dd tier's value
    exact = getattr(oracle, f"{func}_q")(A)
    assert np.all(np.isfinite(bound))
    assert np.all(np.abs(v - exact) <= bound)
    # the final rounding to binary64 dominates unless sum|term| / |value| is huge
    scale = getattr(oracle, f"{func}_abs")(A)
    assert np.all(bound <= 4 * EPS * (np.abs(v.real) + np.abs(v.imag)) + 1e-25 * scale)


@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_dd_certified_large(ext, oracle, func):
    print(f"{func} dd certified N=16", flush=True)
    A = rng_symmetric(2, 16, 80 + (func == "lhaf"), 0.5)
    v, bound = getattr(ext, f"{func}_dd_certified")(A)
    assert np.array_equal(v, getattr(ext, f"{func}_dd")(A))
    assert np.all(np.abs(v - getattr(oracle, f"{func}_dd")(A)) <= bound)
    assert np.all(bound <= 8 * EPS * (np.abs(v.real) + np.abs(v.imag)))


def test_dd_certified_cancellation(ext, oracle):
    """[[1, 3], [3, d - 3]] (+) R: lhaf = d lhaf(R) cancels; dd certifies it."""
    g = np.random.default_rng(3)
    R = g.standard_normal((4, 4)) + 1j * g.standard_normal((4, 4))
    A = np.zeros((6, 6), complex)
    A[:2, :2] = [[1.0, 3.0], [3.0, 1e-9 - 3.0]]
    A[2:, 2:] = R + R.T
    v, bound = ext.lhaf_dd_certified(A[None])
    exact = oracle.lhaf_q(A[None])[0]
    assert abs(v[0] - exact) <= bound[0]
    assert bound[0] / abs(exact) < 1e-14
    v0, b0 = ext.haf_dd_certified(np.zeros((2, 0, 0), complex))
    assert np.all(v0 == 1) and np.all(b0 == 0)


@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_fused_small_kernel_equals_per_size_launches(ext, func):
    """The fused small-size kernel (all subset sizes and the final sum in one
    launch, several instances per workgroup) returns the per-subset-size
    launches' values, kappa sums and certified bounds bit for bit, at every
    side it covers, for every workgroup packing."""
    import ctypes
    from gbskernels_amd import _lib
    L = _lib.lib()
    m0, t0 = ctypes.c_int(), ctypes.c_int()
    _lib.check(L.gbs_haf_get_schedule(ctypes.byref(m0), ctypes.byref(t0)), "get_schedule")
    try:
        for N in (2, 4, 6, 8, 10, 12):
            A = rng_symmetric(37, N, 700 + N)
            _lib.check(L.gbs_haf_set_schedule(0, t0.value), "set_schedule")
            ref = [getattr(ext, func)(A), *getattr(ext, f"{func}_kappa")(A), getattr(ext, f"{func}_dd")(A),
                   *getattr(ext, f"{func}_certified")(A), *getattr(ext, f"{func}_dd_certified")(A)]
            for target in (0, 1, 12, 18, 2048):  # the cost model; 4, 3, 2, 1 instances per workgroup
                _lib.check(L.gbs_haf_set_schedule(12, target), "set_schedule")
                got = [getattr(ext, func)(A), *getattr(ext, f"{func}_kappa")(A), getattr(ext, f"{func}_dd")(A),
                       *getattr(ext, f"{func}_certified")(A), *getattr(ext, f"{func}_dd_certified")(A)]
                for k, (a, b) in enumerate(zip(ref, got)):
                    assert np.array_equal(a, b), (N, target, k)
    finally:
        _lib.check(L.gbs_haf_set_schedule(m0.value, t0.value), "set_schedule")


@pytest.mark.parametrize("func", ["haf", "lhaf"])
def test_fused_small_kernel_real_workgroups(ext, func):
    """A workgroup of the fused kernel whose staged matrices are all real runs
    in real arithmetic. Its fp64 values and kappa sums equal the per-size
    launches' (complex arithmetic) bit for bit, imaginary parts +0, at every
    side and packing, in a batch where one instance is complex (its workgroup
    keeps complex arithmetic); -0.0 imaginary parts count as real (values equal)."""
    import ctypes
    from gbskernels_amd import _lib
    L = _lib.lib()
    m0, t0 = ctypes.c_int(), ctypes.c_int()
    _lib.check(L.gbs_haf_get_schedule(ctypes.byref(m0), ctypes.byref(t0)), "get_schedule")

    def bits(x):
        return np.ascontiguousarray(x).view(np.uint64)

    try:
        for N in (2, 4, 6, 8, 10, 12):
            g = np.random.default_rng(900 + N)
            R = g.standard_normal((41, N, N))
            A = (R + R.transpose(0, 2, 1)).astype(np.complex128)
            C = g.standard_normal((N, N))
            A[17] += 1e-3j * (C + C.T)
            Aneg = np.zeros((8, N, N), np.complex128)
            Aneg.real = A[:8].real
            Aneg.imag = -0.0
            _lib.check(L.gbs_haf_set_schedule(0, t0.value), "set_schedule")
            ref = [getattr(ext, func)(A), *getattr(ext, f"{func}_kappa")(A)]
            ref_neg = getattr(ext, func)(Aneg)
            for target in (0, 1, 12, 18, 2048):  # the cost model; 4, 3, 2, 1 instances per workgroup
                _lib.check(L.gbs_haf_set_schedule(12, target), "set_schedule")
                got = [getattr(ext, func)(A), *getattr(ext, f"{func}_kappa")(A)]
                for a, b in zip(ref, got):
                    assert np.array_equal(bits(a), bits(b)), (N, target)
                real = np.arange(41) != 17
                assert np.all(got[0][real].imag == 0)
                assert np.array_equal(getattr(ext, func)(Aneg), ref_neg), (N, target)
    finally:
        _lib.check(L.gbs_haf_set_schedule(m0.value, t0.value), "set_schedule")
