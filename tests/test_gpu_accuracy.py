"""Accuracy against the mpmath ref tier (80 digits), the bar the backend is held
to: fp64 within 1e-12 relative on well-conditioned physical inputs, double-
double recovering full binary64 accuracy under cancellation, the auto tier
escalating exactly the cancelling cases, and every certified bound (fp64 and
dd) enclosing the mpmath value. The error is taken in mpmath."""
import mpmath as mp
import numpy as np
import pytest

import gbskernels_amd as gk
from gbskernels_amd import _ref
from gbskernels_amd.workloads import cancellation, gbs_adjacency, haar_unitary, haf_subset, perm_rows, threshold_O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from gbskernels_amd.ext import gbskernels_ext
    return gbskernels_ext


def physical(func, n, K, seed):
    rng = np.random.default_rng(seed)
    if func == "perm":
        U = haar_unitary(64, rng)
        return np.stack([perm_rows(U, n, rng) for _ in range(K)])
    if func in ("haf", "lhaf"):
        A = gbs_adjacency(rng)
        return np.stack([haf_subset(A, n, rng, func == "lhaf") for _ in range(K)])
    return np.stack([threshold_O(n, rng) for _ in range(K)]).astype(np.complex128)


def ext_kappa(ext, func, mats):
    v, absn = getattr(ext, f"{func}_kappa")(mats)
    return absn / np.abs(v)


def rel(vals, exact):
    with mp.workdps(80):
        return np.array([float(abs(mp.mpc(complex(v)) - e) / abs(e)) for v, e in zip(vals, exact)])


def encloses(vals, bounds, exact):
    with mp.workdps(80):
        return all(np.isfinite(b) and abs(mp.mpc(complex(v)) - e) <= mp.mpf(float(b))
                   for v, b, e in zip(vals, bounds, exact))


@pytest.mark.parametrize("func,n", [("perm", 8), ("perm", 12), ("haf", 8), ("haf", 12), ("lhaf", 8),
                                    ("lhaf", 10), ("tor", 4), ("tor", 6)])
def test_physical_every_tier_vs_mpmath(ext, func, n):
    print(f"accuracy {func} n={n}", flush=True)
    mats = physical(func, n, 3, 50 + n)
    exact = [_ref.REF[func](m, dps=80) for m in mats]
    f = getattr(gk, f"{func}_batched")
    # 1e-12 relative on well-conditioned elements (kappa = sum|term| / |value|
    # <= 1e3); the torontonians here reach kappa 7e5, where binary64 rounding
    # alone gives ~1e-11: those are held to the kappa-scaled error 64 n u kappa
    kappa = ext_kappa(ext, func, mats)
    well = kappa <= 1e3
    bar = np.where(well, 1e-12, 64 * n * 2.0 ** -53 * kappa)
    assert np.all(rel(f(mats, precision="fp64", backend="gpu"), exact) <= bar), kappa
    assert rel(f(mats, precision="dd", backend="gpu"), exact).max() <= 4e-16
    v, diags = f(mats, precision="auto", backend="gpu", return_diagnostics=True)
    assert np.all(rel(v, exact) <= bar) and all(d["tier"] == "fp64" for d in diags)
    # informative certificates: fp64 bounds carry the rounding depth times the
    # cancellation (sum|term| / |value| reaches 1e3 for these torontonians); dd
    # bounds are dominated by the final rounding to binary64
    for name, width in (("certified", 1e-6), ("dd_certified", 1e-15)):
        v, b = getattr(ext, f"{func}_{name}")(mats)
        assert encloses(v, b, exact), name
        assert max(float(mp.mpf(float(bb)) / abs(e)) for bb, e in zip(b, exact)) <= width, name


@pytest.mark.parametrize("func", ["perm", "haf", "lhaf", "tor"])
def test_cancelling_dd_and_auto_vs_mpmath(ext, func):
    rng = np.random.default_rng(77)
    mats = np.stack([cancellation(func, d, rng) for d in (1e-2, 1e-10, 1e-14)])
    exact = [_ref.REF[func](m, dps=80) for m in mats]
    f = getattr(gk, f"{func}_batched")
    # the dd tier collapses to binary64: a few ulps whatever the cancellation
    assert rel(f(mats, precision="dd", backend="gpu"), exact).max() <= 4e-16
    v, diags = f(mats, precision="auto", backend="gpu", return_diagnostics=True)
    assert diags[0]["tier"] == "fp64" and diags[1]["tier"] != "fp64" and diags[2]["tier"] != "fp64"
    assert rel(v, exact).max() <= 1e-12
    for name in ("certified", "dd_certified"):
        v, b = getattr(ext, f"{func}_{name}")(mats)
        assert encloses(v, b, exact), name
    # the dd certificate stays informative at delta = 1e-14
    v, b = getattr(ext, f"{func}_dd_certified")(mats)
    assert float(b[2]) <= 1e-12 * abs(complex(exact[2]))
