"""Public-API behaviour on the GPU (reference gbskernels/__init__.py semantics):
precision tiers, the auto tier's per-element rerun, the certified ladder,
Workspace bucketing and buffer reuse, odd loop hafnians."""
import numpy as np
import pytest

import gbskernels_amd as gk
from conftest import rel_err, rng_complex, rng_symmetric
from physics import threshold_O

pytestmark = pytest.mark.gpu


# cancellation families with exact small values (constructions of these tests)
def perm_cancel(delta, seed, n=6):
    """[[1, 2], [-1, 2 + delta]] (+) R: perm = delta * perm(R), Glynn terms O(1)."""
    A = np.zeros((n, n), complex)
    A[:2, :2] = [[1.0, 2.0], [-1.0, 2.0 + delta]]
    A[2:, 2:] = rng_complex((n - 2, n - 2), seed)
    return A


def haf_cancel(delta, seed):
    """4x4 block with matchings 2*2 + 1*1 + 1*(delta - 5) = delta, (+) symmetric R."""
    a = f = 2.0
    b = e = c = 1.0
    d = delta - 5.0
    Bk = np.array([[0, a, b, c], [a, 0, d, e], [b, d, 0, f], [c, e, f, 0]], complex)
    R = rng_symmetric(1, 4, seed)[0]
    np.fill_diagonal(R, 0.0)
    A = np.zeros((8, 8), complex)
    A[:4, :4], A[4:, 4:] = Bk, R
    return A


def lhaf_cancel(delta, seed):
    """[[1, 3], [3, delta - 3]]: lhaf = 3 + (delta - 3) = delta, (+) symmetric R."""
    A = np.zeros((6, 6), complex)
    A[:2, :2] = [[1.0, 3.0], [3.0, delta - 3.0]]
    A[2:, 2:] = rng_symmetric(1, 4, seed)[0]
    return A


def tor_cancel(a):
    return np.array([[a, 0.0], [0.0, a]], complex)


MK = {"perm": lambda d: perm_cancel(d, 3), "haf": lambda d: haf_cancel(d, 4),
      "lhaf": lambda d: lhaf_cancel(d, 5), "tor": tor_cancel}


@pytest.mark.parametrize("func", ["perm", "haf", "lhaf", "tor"])
def test_auto_tier_reruns_only_risky(oracle, func):
    mk = MK[func]
    mats = np.stack([mk(1e-1), mk(1e-10), mk(1e-1)])
    out, diags = getattr(gk, f"{func}_batched")(mats, precision="auto", backend="gpu", return_diagnostics=True)
    tiers = [d["tier"] for d in diags]
    assert tiers[0] == "fp64" and tiers[2] == "fp64"
    assert tiers[1] in ("dd", "ref")
    assert diags[1]["cancellation"] >= gk._AUTO_KAPPA_MAX
    exact = oracle.QUAD[func](mats)
    assert rel_err(out[1], exact[1]) < 1e-12
    fp64 = getattr(gk, f"{func}_batched")(mats[1:2], backend="gpu")[0]
    assert rel_err(fp64, exact[1]) > 1e3 * max(rel_err(out[1], exact[1]), 1e-30)
    # well-conditioned elements are the plain fp64 values, bit for bit
    plain = getattr(gk, f"{func}_batched")(mats, backend="gpu")
    assert out[0] == plain[0] and out[2] == plain[2]


@pytest.mark.parametrize("func", ["perm", "haf", "lhaf", "tor"])
def test_certified_api_and_ladder(oracle, func):
    mats = np.stack([MK[func](1e-1), MK[func](1e-9)])
    out, diags = getattr(gk, f"{func}_batched")(mats, precision="certified", backend="gpu", return_diagnostics=True)
    exact = oracle.QUAD[func](mats)
    for i in range(2):
        assert abs(out[i] - exact[i]) <= diags[i]["abs_error_bound"]
        assert diags[i]["tier"] == "certified-fp64"
    # a tight rtol the fp64 certificate cannot prove escalates (here to ref)
    out2, diags2 = getattr(gk, f"{func}_batched")(mats, precision="certified", backend="gpu",
                                                   return_diagnostics=True, rtol=1e-13)
    assert diags2[1]["escalated"] is True and diags2[1]["tier"] in ("certified-dd", "ref")
    assert rel_err(out2[1], exact[1]) < 1e-13
    # single-element form
    v, d = getattr(gk, func)(mats[0], precision="certified", backend="gpu", return_diagnostics=True)
    assert v == out[0] and d["abs_error_bound"] == diags[0]["abs_error_bound"]


@pytest.mark.parametrize("func", ["perm", "haf", "lhaf", "tor"])
def test_certified_ladder_batch_equals_per_element(func):
    """The ladder escalates a whole batch at once (one certified-dd launch for
    every element the fp64 bounds cannot prove); each element's value and
    diagnostics equal those of the same element certified alone."""
    mats = np.stack([MK[func](d) for d in (1e-1, 1e-9, 1e-3, 1e-12, 0.5, 1e-6)])
    out, diags = getattr(gk, f"{func}_batched")(mats, precision="certified", backend="gpu",
                                                 return_diagnostics=True, rtol=1e-10)
    assert any(d["escalated"] for d in diags) and not all(d["escalated"] for d in diags)
    for i in range(len(mats)):
        v, d = getattr(gk, func)(mats[i], precision="certified", backend="gpu", return_diagnostics=True, rtol=1e-10)
        assert v == out[i] and d == diags[i], (i, d, diags[i])


def test_certified_refuses_nonfinite_input():
    A = np.eye(4)
    A[0, 0] = np.nan
    with pytest.raises(ValueError, match="finite"):
        gk.perm_batched(A[None], precision="certified", return_diagnostics=True)


def test_odd_loop_hafnian_is_augmented(oracle):
    A = rng_symmetric(5, 7, 77)
    got = gk.lhaf_batched(A, backend="gpu")
    assert rel_err(got, oracle.matchings(A, loop=True)).max() < 1e-11
    got_dd = gk.lhaf_batched(A, precision="dd", backend="gpu")
    assert rel_err(got_dd, oracle.matchings(A, loop=True)).max() < 1e-12


def test_workspace_bucketing_and_reuse():
    mats = [rng_complex((n, n), 10 + i) for i, n in enumerate([5, 9, 5, 12, 9, 3])]
    with gk.Workspace(backend="gpu") as ws:
        r1 = ws.reallocs()
        small = ws.perm_batched(mats[:3])
        r2 = ws.reallocs()
        res = ws.perm_batched(mats)
        r3 = ws.reallocs()
        again = ws.perm_batched(mats[:3])
        r4 = ws.reallocs()
    alone = np.array([gk.perm(m, backend="gpu") for m in mats])
    assert np.array_equal(res, alone)                 # bucketing never changes values
    assert np.array_equal(small, again) and np.array_equal(small, alone[:3])
    assert r1 == 0 and r2 >= 1 and r3 >= r2 and r4 == r3   # a smaller call does not grow


def test_workspace_ragged_mixed_sizes_every_function():
    """One ragged call per function (sizes interleaved, odd loop hafnians, the
    empty matrix): every value bit-identical to the uniform call on that matrix
    alone; an oversize member is rejected naming its index."""
    g = np.random.default_rng(41)
    cases = {
        "perm": [rng_complex((n, n), 60 + i) for i, n in enumerate([3, 14, 1, 9, 14, 0, 20])],
        "haf": list(rng_symmetric(1, 4, 61)) + list(rng_symmetric(2, 10, 62)) + list(rng_symmetric(1, 16, 63)),
        "lhaf": list(rng_symmetric(1, 5, 64)) + list(rng_symmetric(2, 8, 65)) + list(rng_symmetric(1, 7, 66)),
        "tor": [threshold_O(nm, g).astype(complex) for nm in (2, 5, 2, 7, 3)],
    }
    with gk.Workspace(backend="gpu") as ws:
        for func, mats in cases.items():
            got = getattr(ws, f"{func}_batched")(mats)
            alone = np.array([getattr(gk, func)(m, backend="gpu") for m in mats])
            assert np.array_equal(got, alone), func
        with pytest.raises(ValueError, match="matrix #1 "):
            ws.perm_batched([rng_complex((4, 4), 1), rng_complex((30, 30), 2)])
    with pytest.raises(RuntimeError):
        ws.perm_batched(cases["perm"])


def test_workspace_auto_equals_alone():
    mats = [haf_cancel(1e-10, 1), rng_symmetric(1, 6, 2)[0], haf_cancel(1e-1, 3)]
    with gk.Workspace(backend="gpu") as ws:
        res = ws.haf_batched(mats, precision="auto")
    alone = np.array([gk.haf(m, precision="auto", backend="gpu") for m in mats])
    assert np.array_equal(res, alone)


def test_workspace_resident():
    A = rng_complex((30, 10, 10), 8)
    with gk.Workspace(backend="gpu") as ws:
        h = ws.perm_resident(A)
        assert np.array_equal(h.numpy(), gk.perm_batched(A, backend="gpu"))


def test_resident_dlpack_into_torch():
    """The resident result goes to PyTorch without a host round trip
    (DLPack, device type ROCm); the consumer owns and frees the buffer."""
    import gc
    import torch
    from gbskernels_amd.ext import gbskernels_ext as ext
    A = rng_complex((40, 9, 9), 12)
    want = gk.perm_batched(A, backend="gpu")
    with gk.Workspace(backend="gpu") as ws:
        h = ws.perm_resident(A)
        assert h.__dlpack_device__() == (ext.KDL_ROCM, 0)
        t = torch.from_dlpack(h)
        assert t.is_cuda and t.dtype == torch.complex128 and tuple(t.shape) == (40,)
        assert np.array_equal(t.cpu().numpy(), want)
        assert h.ptr is None                      # ownership moved to the tensor
        with pytest.raises(BufferError):
            h.__dlpack__()
        n_live = len(ext._exported)
        del t
        gc.collect()
        torch.cuda.synchronize()
        assert len(ext._exported) == n_live - 1   # the consumer called the deleter
        cap = ws.haf_resident(rng_symmetric(3, 6, 2)).__dlpack__()
        del cap                                   # never consumed: freed by the capsule
        assert len(ext._exported) == n_live - 1


def test_tor_api_real_domain(oracle):
    from physics import threshold_O
    rng = np.random.default_rng(3)
    O = np.stack([threshold_O(6, rng) for _ in range(4)])
    got = gk.tor_batched(O, backend="gpu")
    exact = oracle.tor_dd(O.astype(complex))
    assert np.all(np.abs(got - exact) <= 64 * 6 * 2.0 ** -52 * oracle.tor_abs(O.astype(complex)))
