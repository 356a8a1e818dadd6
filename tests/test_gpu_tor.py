"""GPU parity of the real-domain torontonian kernel (fp64, kappa) against the
oracle (DD subset determinants, the reference's recursive Cholesky), closed
forms, off-domain refusal and the reference's published checksums."""
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_err
from physics import threshold_O

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu
EPS = 2.0 ** -52


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


def tor_bound(abs_sum, nm):
    return 64.0 * max(nm, 1) * EPS * abs_sum


def physical_stack(nm, B, seed):
    rng = np.random.default_rng(seed)
    return np.stack([threshold_O(nm, rng) for _ in range(B)]).astype(np.complex128)


@pytest.mark.parametrize("nm", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_tor_physical_vs_dd_oracle(ext, oracle, nm):
    print(f"tor fp64 nm={nm}", flush=True)
    O = physical_stack(nm, 6, 100 + nm)
    got = ext.tor(O)
    exact = oracle.tor_dd(O)
    assert np.all(np.abs(got - exact) <= tor_bound(oracle.tor_abs(O), nm))
    assert np.all(got.imag == 0)


@pytest.mark.parametrize("nm", [12, 16, 20])
def test_tor_large_vs_recursive_oracle(ext, oracle, nm):
    print(f"tor fp64 large nm={nm}", flush=True)
    O = physical_stack(nm, 2, 200 + nm)
    got = ext.tor(O).real
    ref = oracle.tor_recursive(O.real)
    # sum|term| of the real domain = sum 1/sqrt(det) is computed on the GPU (kappa)
    _, absn = ext.tor_kappa(O)
    assert np.all(np.abs(got - ref) <= 2 * tor_bound(absn, nm))


@pytest.mark.parametrize("nm", [1, 3, 6, 9, 13])
def test_tor_closed_form(ext, nm):
    a = 0.2
    got = ext.tor((a * np.eye(2 * nm))[None])[0]
    assert got.real == pytest.approx((a / (1 - a)) ** nm, rel=1e-12)
    assert ext.tor(np.zeros((1, 2 * nm, 2 * nm)))[0] == pytest.approx(0.0, abs=1e-15)


def test_tor_zero_modes(ext):
    assert ext.tor(np.zeros((3, 0, 0)))[0] == 1.0


def test_tor_off_domain_falls_back_to_complex(ext, oracle):
    """Real O with I - O_S not SPD: the real kernel refuses (non-positive pivot)
    and the complex-symmetric kernel recomputes the instance (the reference's
    recursive -> complex-LU fallback, core/host_api.cu:177-228)."""
    O = np.stack([1.5 * np.eye(8), 0.2 * np.eye(8)]).astype(complex)
    got = ext.tor(O)
    assert got[0] == pytest.approx(oracle.tor(O[:1])[0], rel=1e-12)   # = (2 - 1)^4 = 1
    assert got[0].real == pytest.approx(1.0, rel=1e-12)
    assert got[1].real == pytest.approx(0.25 ** 4, rel=1e-12)
    # mixed-sign determinants take the principal branch
    X = np.diag([1.5, 0.3, 0.3, 0.3]).astype(complex)  # mode 0: det(I - O) = -0.35
    assert ext.tor(X[None])[0] == pytest.approx(oracle.tor(X[None])[0], rel=1e-12)


@pytest.mark.parametrize("nm", [1, 3, 5, 7, 9])
def test_tor_complex_symmetric_vs_oracle(ext, oracle, nm):
    print(f"tor complex nm={nm}", flush=True)
    g = np.random.default_rng(800 + nm)
    X = 0.08 * (g.standard_normal((4, 2 * nm, 2 * nm)) + 1j * g.standard_normal((4, 2 * nm, 2 * nm)))
    O = X + X.transpose(0, 2, 1)
    got = ext.tor(O)
    want, absn = oracle.tor(O, want_abs=True)
    assert np.all(np.abs(got - want) <= 2 * tor_bound(absn, nm))
    v, a = ext.tor_kappa(O)
    assert np.array_equal(v, got)
    assert rel_err(a, absn).max() < 1e-10


@pytest.mark.parametrize("nm", [1, 2, 4, 6, 8])
def test_tor_hermitian_and_general_complex_vs_oracle(ext, oracle, nm):
    """Complex O that is not symmetric -- the complex-basis O = I - Q^-1 of a
    physical state (Hermitian) and a general perturbation of it -- takes the LU
    fallback and agrees with the oracle's partial-pivot LU (principal branch);
    the Hermitian one's value is the real-basis torontonian's."""
    from gbskernels_amd.gbs import threshold_O_xxpp
    from gbskernels_amd.sampling import qmat
    print(f"tor hermitian nm={nm}", flush=True)
    g = np.random.default_rng(900 + nm)
    r = g.uniform(0.2, 0.6, nm)
    z = (g.standard_normal((nm, nm)) + 1j * g.standard_normal((nm, nm))) / np.sqrt(2.0)
    U, rr = np.linalg.qr(z)
    U = U * (np.diagonal(rr) / np.abs(np.diagonal(rr)))
    S = np.block([[U.real, -U.imag], [U.imag, U.real]]) @ np.diag(np.concatenate([np.exp(-r), np.exp(r)]))
    cov = 0.7 * (S @ S.T) + 0.3 * np.eye(2 * nm)      # lossy: a mixed state
    O = np.eye(2 * nm) - np.linalg.inv(qmat(cov))
    assert not np.allclose(O, O.T)
    Y = O + 0.02 * (g.standard_normal(O.shape) + 1j * g.standard_normal(O.shape))
    stack = np.stack([O, Y])
    got = ext.tor(stack)
    want, absn = oracle.tor(stack, want_abs=True)
    assert np.all(np.abs(got - want) <= 2 * tor_bound(absn, nm))
    Ox, _ = threshold_O_xxpp(cov)
    assert abs(got[0] - ext.tor_single(Ox, 0)) <= 4 * tor_bound(absn[:1], nm)[0]


@pytest.mark.parametrize("nm", [3, 8, 12])
def test_tor_kappa_value_bitwise_and_abs(ext, oracle, nm):
    O = physical_stack(nm, 8, 300 + nm)
    v, a = ext.tor_kappa(O)
    assert np.array_equal(v, ext.tor(O))
    if nm <= 8:
        assert rel_err(a, oracle.tor_abs(O)).max() < 1e-11


def test_tor_batch_composition_independence(ext):
    O = physical_stack(13, 12, 400)
    full = ext.tor(O)
    for idx in ([0], [5, 2], list(range(11, -1, -2))):
        assert np.array_equal(ext.tor(O[idx]), full[idx])


def test_reference_published_checksums_recursive(ext, oracle, ref_checksums):
    for row in ref_checksums:
        if row["func"] != "tor_recursive":
            continue
        dim, B = row["matrix_dim"], row["batch"]
        print(f"checksum tor_recursive dim={dim}", flush=True)
        a = oracle.gen_bench_real(dim, B, 0.05).astype(np.complex128)
        got = float(np.sum(np.abs(ext.tor(a).real)))
        # the fp64 cells cancel: the reference's own two real kernels differ by
        # 3e-8 relative at dim 16; our value is checked against the DD oracle
        want_dd = float(np.sum(np.abs(oracle.tor_dd(a[:256]).real)))
        got256 = float(np.sum(np.abs(ext.tor(a[:256]).real)))
        assert got256 == pytest.approx(want_dd, rel=1e-9 if dim < 16 else 1e-6)
        assert got == pytest.approx(row["checksum"], rel=1e-9 if dim < 16 else 1e-7), dim


# ------------------------------------------------------------------- dd tier
@pytest.mark.parametrize("nm", [1, 2, 4, 6, 8, 10, 12])
def test_tor_dd_vs_dd_oracle(ext, oracle, nm):
    print(f"tor dd nm={nm}", flush=True)
    O = physical_stack(nm, 4, 500 + nm)
    got = ext.tor_dd(O)
    exact = oracle.tor_dd(O)  # DD subset determinants (core/torontonian_dd.cu algorithm)
    # both are double-double evaluations of an alternating sum: agree to ~u^2 * sum|term|
    assert np.all(np.abs(got - exact) <= 1e-13 * np.abs(exact) + 1e-27 * oracle.tor_abs(O))


@pytest.mark.parametrize("nm", [14, 20])
def test_tor_dd_large_vs_fp64(ext, oracle, nm):
    O = physical_stack(nm, 2, 600 + nm)
    got = ext.tor_dd(O).real
    fp = ext.tor(O).real
    _, absn = ext.tor_kappa(O)
    assert np.all(np.abs(got - fp) <= tor_bound(absn, nm))


def test_tor_dd_closed_form_cancellation(ext):
    # single mode, O = a I: tor = a / (1 - a); fp64 cancels as a -> 0, dd does not
    for a in (1e-3, 1e-8, 1e-12):
        O = np.array([[[a, 0.0], [0.0, a]]], complex)
        want = a / (1 - a)
        assert ext.tor_dd(O)[0].real == pytest.approx(want, rel=1e-15)
    assert np.isnan(ext.tor_dd(np.array([1.5 * np.eye(4)], complex))[0].real)


def test_tor_fallback_scan_large_sparse(ext):
    """A batch with more partial slots than one pass of the fallback scan covers
    (20 modes: 512 slots per instance, 1100 instances) and three complex
    instances spread through it: the fallback recomputes exactly those, bit for
    bit as when each is evaluated alone, and leaves the real ones untouched."""
    print("tor fallback scan nm=20 B=1100", flush=True)
    rng = np.random.default_rng(77)
    base = np.stack([threshold_O(20, rng) for _ in range(1100)]).astype(np.complex128)
    stack = base.copy()
    idx = [0, 600, 1099]
    for i in idx:
        stack[i, 0, 1] += 1e-3j
        stack[i, 1, 0] += 1e-3j
    got = ext.tor(stack)
    real_only = ext.tor(base)
    keep = np.ones(len(base), bool)
    keep[idx] = False
    assert np.array_equal(got[keep], real_only[keep])
    for i in idx:
        assert np.isfinite(got[i]) and got[i].imag != 0.0
        assert np.array_equal(got[i:i + 1], ext.tor(stack[i:i + 1]))


def test_tor_dd_reads_the_real_part_under_imaginary_noise(oracle):
    """A real O carrying 1e-15 imaginary noise (below the package's 1e-12 gate)
    evaluates at precision 'dd' and 'auto' as its real part does, as the
    reference's dd kernel (which reads only the real part) does; the C ABI's
    tor_dd gives the same value as on the noise-free input."""
    import gbskernels_amd as gk
    import gbskernels_ext
    rng = np.random.default_rng(5150)
    nm = 6
    O = physical_stack(nm, 5, 5151)
    noisy = O + 1e-15j * rng.standard_normal(O.shape)
    exact = oracle.tor_dd(O)
    assert np.array_equal(gbskernels_ext.tor_dd(noisy), gbskernels_ext.tor_dd(O))
    got_dd = gk.tor_batched(noisy, precision="dd", backend="gpu")
    assert np.all(np.isfinite(got_dd))
    assert np.all(np.abs(got_dd - exact) <= 1e-13 * np.abs(exact) + 1e-27 * oracle.tor_abs(O))
    got_auto, diags = gk.tor_batched(noisy, precision="auto", backend="gpu", return_diagnostics=True)
    assert np.all(np.isfinite(got_auto))
    # the fp64 pass sees the complex input (fallback path): the noise itself
    # moves the value by about 1e-15 times the conditioning of I - O_S
    assert np.all(np.abs(got_auto - exact) <= 1e-10 * oracle.tor_abs(O))


def test_tor_schedules_bfs_identical_grouped_correct(ext, oracle):
    """The breadth-first middle walk returns the depth-first walk's values and
    auxiliaries bit for bit (every tier); the grouped small-count kernel
    (several instances per wave, trailing modes first: a different split)
    agrees with the oracle, keeps the fp64 / kappa / certified values identical
    to each other, and its certificates enclose."""
    import ctypes
    from gbskernels_amd import _lib
    L = _lib.lib()
    saved = []
    for fam in (0, 1):
        g, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(L.gbs_tor_get_schedule(fam, ctypes.byref(g), ctypes.byref(b)), "get_schedule")
        saved.append((g.value, b.value))

    def run_all(O):
        return (ext.tor(O), *ext.tor_kappa(O), *ext.tor_certified(O), ext.tor_dd(O), *ext.tor_dd_certified(O))
    try:
        for nm in (3, 6, 8, 10, 12):
            print(f"tor schedules nm={nm}", flush=True)
            O = physical_stack(nm, 5, 800 + nm)
            res = {}
            for label, gm, bfs in (("dfs", 0, 0), ("bfs", 0, 32), ("grouped", 10, 32)):
                for fam in (0, 1):
                    _lib.check(L.gbs_tor_set_schedule(fam, gm, bfs), "set_schedule")
                res[label] = run_all(O)
            for a, b in zip(res["dfs"], res["bfs"]):
                assert np.array_equal(a, b)
            v, kv, ka, cv, cb, dv, ddv, ddb = res["grouped"]
            assert np.array_equal(v, kv) and np.array_equal(v, cv) and np.array_equal(dv, ddv)
            exact = oracle.tor_dd(O)
            absn = oracle.tor_abs(O)
            assert np.all(np.abs(v - exact) <= tor_bound(absn, nm))
            assert rel_err(ka, absn).max() < 1e-12
            assert np.all(np.abs(cv - exact) <= cb)
            assert np.all(np.abs(dv - exact) <= 1e-13 * np.abs(exact) + 1e-27 * absn)
            assert np.all(np.abs(ddv - exact) <= ddb)
    finally:
        for fam, (g, b) in enumerate(saved):
            _lib.check(L.gbs_tor_set_schedule(fam, g, b), "set_schedule")
