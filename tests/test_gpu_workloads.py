"""GPU parity on the benchmark workloads themselves, at their real sizes:

* the permanent workload's 24x24 Haar-row matrices (the dominant kernel)
  against the oracle's double-double Glynn walk, and the dd tier on its 16x16
  Haar rows;
* the certified torontonian of lossy two-mode-squeezed states at 12, 16 and 20
  modes: the certified-fp64 interval must contain the certified-dd interval
  (nested intervals: a one-sided fp64 bound error would show), and certified-dd
  at 20 modes against the oracle's double-double subset determinants;
* the served sampler stream (perm / haf / lhaf / tor, sizes geometric(0.15) on
  2..28, tiers 70% fp64 / 25% dd / 5% certified, Poisson arrivals at 50k req/s)
  replayed through the continuous-batching server: every value and auxiliary
  bit-identical to the item evaluated alone, every certificate finite and
  enclosing the dd value, and the largest sizes against the dd tier and the
  oracle."""
import sys
from pathlib import Path

import numpy as np
import pytest

from physics import threshold_O

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "gbskernels_amd" / "ext"))
pytestmark = pytest.mark.gpu
U = 2.0 ** -53


@pytest.fixture(scope="module")
def ext():
    import gbskernels_ext
    return gbskernels_ext


# ------------------------------------------------------------- permanent
def test_perm_n24_haar_rows_vs_oracle_dd(ext, oracle):
    print("perm workload n=24 Haar rows vs oracle dd", flush=True)
    _, m24 = __import__("gbskernels_amd.workloads", fromlist=["x"]).perm_haar_rows_workload(take16=0, take24=8)
    got, absn = ext.perm_kappa(m24)
    assert np.array_equal(got, ext.perm(m24))
    exact = oracle.perm_dd(m24)
    # the Glynn walk's condition number sum|term| / |perm| is 3e3 - 5e4 on these
    # rows (repeated rows of a unitary cancel), so relative errors of ~1e-12
    # are what binary64 can give: the error is held to the walk's error scale
    # 64 n u kappa per element (the scale every fp64 permanent test uses)
    kappa = absn / np.abs(exact)
    rel = np.abs(got - exact) / np.abs(exact)
    print(f"kappa {kappa.min():.3g}..{kappa.max():.3g}, rel err max {rel.max():.3g}", flush=True)
    assert np.all(rel <= 64 * 24 * U * kappa), (rel, kappa)
    assert np.all(rel[kappa <= 100] <= 1e-12)
    # the dd tier at 24 agrees with the oracle's dd walk to the last binary64 bit
    dd = ext.perm_dd(m24)
    assert np.all(np.abs(dd - exact) <= 2 * U * np.abs(exact))


def test_perm_dd_n16_haar_rows_vs_oracle_dd(ext, oracle):
    from gbskernels_amd.workloads import perm_haar_rows_workload
    m16, _ = perm_haar_rows_workload(take16=256, take24=0)
    got = ext.perm_dd(m16)
    exact = oracle.perm_dd(m16)
    assert np.all(np.abs(got - exact) <= 2 * U * np.abs(exact))
    fp = ext.perm(m16)
    absn = oracle.perm_abs(m16)
    assert np.all(np.abs(fp - exact) <= 64 * 16 * U * absn)


# ------------------------------------------------------------- torontonian
def physical_stack(nm, B, seed):
    rng = np.random.default_rng(seed)
    return np.stack([threshold_O(nm, rng) for _ in range(B)]).astype(np.complex128)


@pytest.mark.parametrize("nm", [12, 16, 20])
def test_tor_certified_nested_in_certified_dd(ext, nm):
    print(f"tor certified nested intervals nm={nm}", flush=True)
    O = physical_stack(nm, 4 if nm < 20 else 2, 700 + nm)
    v, b = ext.tor_certified(O)
    vd, bd = ext.tor_dd_certified(O)
    assert np.all(np.isfinite(b)) and np.all(np.isfinite(bd))
    assert np.all(bd < b)
    # [vd - bd, vd + bd] (which holds the exact value) lies inside [v - b, v + b]
    assert np.all(np.abs(v - vd) <= b - bd), (np.abs(v - vd), b, bd)
    assert np.array_equal(v, ext.tor(O)) and np.array_equal(vd, ext.tor_dd(O))


def test_tor_dd_certified_nm20_vs_oracle_dd(ext, oracle):
    print("tor certified-dd nm=20 vs oracle dd", flush=True)
    O = physical_stack(20, 2, 720)
    v, b = ext.tor_dd_certified(O)
    ref = oracle.tor_dd(O)            # DD subset determinants: error far below one binary64 ulp
    assert np.all(np.isfinite(b)) and np.all(np.abs(v - ref) <= b)
    # informative: at 20 modes the dd sum's own error (2^20 terms, cancelling)
    # brings the bound to about ten binary64 ulps of the value
    assert np.all(b <= 64 * 2.0 ** -52 * np.abs(v))


# ------------------------------------------------------------- served stream
def _alone(ext, func, tier, A):
    from gbskernels_amd.serving import _as_matrix
    name = func if tier == "fp64" else f"{func}_{tier}"
    r = getattr(ext, name)(_as_matrix(func, A)[None])
    if isinstance(r, tuple):
        return r[0][0], r[1][0]
    return r[0], 0.0


def test_served_sampler_stream(ext, oracle):
    from gbskernels_amd.serving import Server
    from gbskernels_amd.workloads import served_mix
    items, arrivals = served_mix(2000, seed=55, funcs=("perm", "haf", "lhaf", "tor"), nmax=28,
                                 tiers=(("fp64", 0.70), ("dd", 0.25), ("certified", 0.05)))
    sizes = np.array([A.shape[0] for _, _, A in items])
    print(f"served stream: {len(items)} requests, max dim {sizes.max()}", flush=True)
    with Server(slots=4, streams_per_slot=4, max_batch=512) as s:
        vals, aux, infos = s.replay(items, arrivals)
    assert all(i.status == 0 for i in infos)
    big_checked = {"perm": 0, "haf": 0, "lhaf": 0}
    for i, ((f, t, A), v, a) in enumerate(zip(items, vals, aux)):
        w, c = _alone(ext, f, t, A)
        assert _bits(v) == _bits(w) and _bits(a) == _bits(c), (i, f, t, A.shape)
        if t == "certified":
            assert np.isfinite(a) and a >= 0, (i, f, A.shape)
            if not (f == "tor" and A.shape[0] > 2 * 20):
                vd = _alone(ext, f, "dd", A)[0]
                assert abs(v - vd) <= a, (i, f, A.shape)
        if f != "tor" and A.shape[0] >= 25 and t == "fp64":
            # the largest sizes: the fp64 value against the dd tier, within the
            # fp64 error scale n u sum|term| (sum|term| from the kappa tier)
            _, absn = _alone(ext, f, "kappa", A)
            vd = _alone(ext, f, "dd", A)[0]
            assert abs(v - vd) <= 64 * A.shape[0] * U * absn, (i, f, A.shape)
            big_checked[f] += 1
    assert big_checked["perm"] + big_checked["haf"] + big_checked["lhaf"] >= 3, big_checked
    # the smallest of the largest permanents and hafnians against the oracle's
    # dd walks: two double-double evaluations agree to a few binary64 ulps
    for f, fn in (("perm", oracle.perm_dd), ("haf", oracle.haf_dd)):
        cand = [A for (g, t, A) in items if g == f and A.shape[0] >= 25]
        if not cand:
            continue
        A = min(cand, key=lambda M: M.shape[0])
        print(f"served stream: {f} n={A.shape[0]} against the oracle's dd walk", flush=True)
        exact = fn(A[None])[0]
        vd = _alone(ext, f, "dd", A)[0]
        absn = _alone(ext, f, "kappa", A)[1]
        assert abs(vd - exact) <= 4 * U * abs(exact) + 1e-28 * absn, (f, A.shape)


def _bits(x):
    return np.array([x], np.complex128).view(np.uint64).tolist()
