"""Heterogeneous batches and the continuous-batching server on the GPU.

The contract checked here: an item's value (and kappa / certificate) is
bit-identical whether it is evaluated alone through the uniform API or inside
any mix of functions, tiers and sizes -- in one batch, replayed from a captured
graph, or served while other requests are on the device. Values themselves are
pinned against the oracle by the per-function test files; here the uniform API
is the reference."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import rng_complex, rng_symmetric
from physics import threshold_O

pytestmark = pytest.mark.gpu

ext = None


def _ext():
    global ext
    if ext is None:
        from gbskernels_amd.ext import gbskernels_ext as e
        ext = e
    return ext


def alone(func, tier, A):
    """Value and aux of one item through the uniform (drop-in) API."""
    from gbskernels_amd.serving import _as_matrix
    name = func if tier == "fp64" else f"{func}_{tier}"
    r = getattr(_ext(), name)(_as_matrix(func, A)[None])
    if isinstance(r, tuple):
        return complex(r[0][0]), float(r[1][0])
    return complex(r[0]), 0.0


def make_item(func, dim, seed):
    if func == "perm":
        return rng_complex((dim, dim), seed, 0.5)
    if func in ("haf", "lhaf"):
        return rng_symmetric(1, dim, seed, 0.4)[0]
    return threshold_O(dim // 2, np.random.default_rng(seed))


TIERS_OF = {"perm": ("fp64", "dd", "kappa", "certified"), "haf": ("fp64", "dd", "kappa", "certified"),
            "lhaf": ("fp64", "dd", "kappa", "certified"), "tor": ("fp64", "dd", "kappa", "certified")}


def mixed_items(n, seed, dims=(0, 1, 2, 3, 4, 6, 8, 11, 12, 16, 20)):
    g = np.random.default_rng(seed)
    items = []
    for i in range(n):
        f = ("perm", "haf", "lhaf", "tor")[g.integers(4)]
        t = TIERS_OF[f][g.integers(len(TIERS_OF[f]))]
        d = int(dims[g.integers(len(dims))])
        if f == "tor":
            d = 2 * max(1, d // 2)
        if f == "haf" and d % 2:
            d += 1
        items.append((f, t, make_item(f, d, seed * 1000 + i)))
    return items


def same(v, a, w, b):
    """bitwise equality of value and aux (NaN == NaN)"""
    vb = np.array([v], np.complex128).view(np.uint64)
    wb = np.array([w], np.complex128).view(np.uint64)
    return np.array_equal(vb, wb) and np.array_equal(np.array([a]).view(np.uint64), np.array([b]).view(np.uint64))


def test_batch_mixed_bitwise_equals_alone():
    from gbskernels_amd.serving import Batch
    items = mixed_items(160, 1)
    with Batch(nstreams=4, use_graphs=False) as b:
        vals, aux = b.run(items)
        st = b.stats()
    assert st["batches"] == 1 and st["items"] == len(items) and st["buckets"] > 10
    for (f, t, A), v, a in zip(items, vals, aux):
        w, c = alone(f, t, A)
        assert same(v, a, w, c), (f, t, A.shape, v, w, a, c)


def test_batch_graph_replays_match_eager():
    from gbskernels_amd.serving import Batch
    dims = (4, 8, 12, 16)
    with Batch(nstreams=3, use_graphs=True) as gb, Batch(nstreams=3, use_graphs=False) as eb:
        for rep in range(6):
            items = mixed_items(48, 100 + rep % 2, dims)  # same signature every call
            # new data, same shapes: the replay must read the freshly packed inputs
            items = [(f, t, A * (1.0 + 0.125 * rep)) for f, t, A in items]
            gv, ga = gb.run(items)
            _ext().perm(rng_complex((7, 9, 9), rep))  # eager work between replays
            ev, ea = eb.run(items)
            assert np.array_equal(gv.view(np.uint64), ev.view(np.uint64)), rep
            assert np.array_equal(ga.view(np.uint64), ea.view(np.uint64)), rep
        st = gb.stats()
    # two signatures alternate; a capture is dropped only when a buffer grows,
    # which the grow-only buffers stop doing after the first two calls
    assert st["graphs_captured"] + st["graph_replays"] == 6 and st["graph_replays"] >= 3, st


def test_batch_budget_and_errors():
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import Batch, pack
    items = [("perm", "fp64", rng_complex((12, 12), 5))] * 8
    with Batch(budget_bytes=1024) as b:
        with pytest.raises(_lib.GbskError):
            b.run(items)
    with Batch() as b:
        # a tier code the library does not know is refused as unimplemented
        f, t, d, flat, off = pack([("perm", "fp64", rng_complex((3, 3), 1))])
        with pytest.raises(NotImplementedError):
            b.run_packed(f, np.full(1, 9, np.int32), d, flat, off)
        with pytest.raises(ValueError):
            b.run([("perm", "fp64", rng_complex((30, 30), 1))])
        # the handle stays usable after a rejected batch
        v, _ = b.run(items[:2])
        assert v[0] == v[1] == alone("perm", "fp64", items[0][2])[0]


def test_server_submit_wait_bitwise():
    from gbskernels_amd.serving import Server
    items = mixed_items(120, 7)
    with Server(slots=3, streams_per_slot=4, max_batch=16) as s:
        tickets = [s.submit(f, A, t) for f, t, A in items]
        res = [s.wait(tk) for tk in tickets]
        st = s.stats()
    assert st["completed"] == len(items) and st["batches"] >= len(items) // 16
    for (f, t, A), (v, a, info) in zip(items, res):
        w, c = alone(f, t, A)
        assert info.status == 0 and info.tier_used == t
        assert info.arrival_ns <= info.admitted_ns <= info.done_ns
        assert same(v, a, w, c), (f, t, A.shape)


def test_server_auto_escalates_to_dd(oracle):
    from gbskernels_amd.serving import Server
    d = 1e-11
    # lhaf: [[1, 3], [3, d - 3]] (+) R, lhaf = d lhaf(R); the direct sum spreads
    # the cancellation over many subset terms, which is what kappa sees (a lone
    # 2 x 2 block cancels inside a single power-trace term)
    L = np.zeros((6, 6), complex)
    L[:2, :2] = [[1.0, 3.0], [3.0, d - 3.0]]
    L[2:, 2:] = rng_symmetric(1, 4, 5)[0]
    cancel = {"perm": np.array([[1, 2], [-1, 2 + d]]), "lhaf": L}
    good = rng_complex((10, 10), 3, 0.5)
    with Server(slots=2) as s:
        t_bad = s.submit("perm", cancel["perm"], "auto")
        t_lbad = s.submit("lhaf", cancel["lhaf"], "auto")
        t_good = s.submit("perm", good, "auto")
        vb, ab, ib = s.wait(t_bad)
        vl, al, il = s.wait(t_lbad)
        vg, ag, ig = s.wait(t_good)
        st = s.stats()
    assert st["escalated"] == 2
    assert ib.tier_used == "dd" and vb == alone("perm", "dd", cancel["perm"])[0]
    assert ab == alone("perm", "kappa", cancel["perm"])[1]
    assert il.tier_used == "dd" and vl == alone("lhaf", "dd", cancel["lhaf"])[0]
    assert ig.tier_used == "kappa" and (vg, ag) == alone("perm", "kappa", good)
    # perm: fl(2 + d) - 2 is exact in binary64 (Sterbenz); lhaf against the oracle
    assert vb == (np.float64(2 + d) - 2.0)
    exact = oracle.lhaf_q(L[None])[0]
    assert abs(vl - exact) <= 1e-12 * abs(exact)


def test_server_replay_poisson_and_threads():
    from gbskernels_amd.serving import Server, latency_percentiles
    g = np.random.default_rng(11)
    items = mixed_items(600, 13, dims=(2, 4, 6, 8, 10, 12, 14, 16))
    arrivals = np.cumsum(g.exponential(1 / 50000.0, len(items)))
    with Server(slots=3, streams_per_slot=4) as s:
        vals, aux, infos = s.replay(items, arrivals)
        # concurrent submitters on the same server
        out = {}

        def worker(k):
            mine = items[k::4]
            ts = [s.submit(f, A, t) for f, t, A in mine]
            out[k] = [s.wait(tk) for tk in ts]
        th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    p = latency_percentiles(infos)
    assert 0 < p["p50"] <= p["p99"] < 30.0
    for i, ((f, t, A), v, a) in enumerate(zip(items, vals, aux)):
        w, c = alone(f, t, A)
        assert same(v, a, w, c), (i, f, t)
        v2, a2, _ = out[i % 4][i // 4]
        assert same(v2, a2, v, a)


def test_server_budget_refuses_oversized_request():
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import Server
    with Server(budget_bytes=4096, slots=2) as s:
        big = s.submit("perm", rng_complex((20, 20), 2), "fp64")   # 6.4 KB of input
        small = s.submit("perm", rng_complex((4, 4), 2), "fp64")
        with pytest.raises(_lib.GbskError):
            s.wait(big)
        v, _, info = s.wait(small)
        assert info.status == 0 and v == alone("perm", "fp64", rng_complex((4, 4), 2))[0]


def test_empty_batches_and_streams():
    from gbskernels_amd.serving import Batch, Server
    with Batch() as b:
        v, a = b.run([])
        assert v.shape == (0,) and a.shape == (0,)
        st = b.stats()
        v1, _ = b.run([("haf", "fp64", np.zeros((0, 0)))])   # haf of the empty matrix
        assert v1[0] == 1
    with Server(slots=1) as s:
        v, a, infos = s.replay([], [])
        assert v.shape == (0,) and infos == []
    assert st["items"] == 0


def test_server_budget_bounds_what_slots_hold():
    """Under a budget the slots' grow-only buffers together never hold more
    than budget_bytes, even when successive batches grow different buffers of
    different slots (inputs for one, workspace for another); values stay
    bit-identical to evaluating each request alone."""
    from gbskernels_amd.serving import Server
    budget = 3 << 20
    g = np.random.default_rng(31)
    waves = [
        [("perm", "fp64", rng_complex((14, 14), 100 + i)) for i in range(400)],     # inputs
        [("haf", "kappa", rng_symmetric(1, 18, 200 + i)[0]) for i in range(60)],    # workspace
        [("perm", "fp64", rng_complex((22, 22), 300 + i)) for i in range(40)],      # workspace
        [("lhaf", "fp64", rng_symmetric(1, 6, 400 + i)[0]) for i in range(900)],    # inputs
    ]
    with Server(budget_bytes=budget, slots=3, streams_per_slot=2) as s:
        for wave in waves:
            order = g.permutation(len(wave))
            tickets = [(i, s.submit(wave[i][0], wave[i][2], wave[i][1])) for i in order]
            for i, tk in tickets:
                v, a, info = s.wait(tk)
                f, t, A = wave[i]
                w, c = alone(f, t, A)
                assert info.status == 0 and same(v, a, w, c), (f, t, A.shape)
        st = s.stats()
    assert 0 < st["device_bytes_max"] <= budget, st


def test_batch_budget_bounds_buffers_across_batches():
    """A budgeted Batch whose successive batches need large inputs, then a large
    workspace, gives its buffers back instead of holding both maxima."""
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import Batch
    big_in = [("perm", "fp64", rng_complex((16, 16), 7 + i)) for i in range(300)]    # 1.2 MB of input
    big_ws = [("haf", "kappa", rng_symmetric(1, 20, 9 + i)[0]) for i in range(20)]
    with Batch(budget_bytes=int(1.6 * (1 << 20))) as b:
        for items in (big_in, big_ws, big_in):
            vals, aux = b.run(items)
            for (f, t, A), v, a in zip(items, vals, aux):
                w, c = alone(f, t, A)
                assert same(v, a, w, c)
