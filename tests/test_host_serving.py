"""Host-side logic of the batch / serving front end and the benchmark
workloads (CPU only, no GPU calls): packing, the odd loop-hafnian augmentation,
the request-stream generator, and the canonical flop counts the roofline uses."""
import math

import numpy as np
import pytest

from conftest import rng_symmetric


def test_pack_offsets_codes_and_augmentation(oracle):
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import pack
    A = rng_symmetric(1, 5, 3)[0]                       # odd loop hafnian
    P = np.arange(16, dtype=np.complex128).reshape(4, 4)
    T = np.eye(6)
    items = [("lhaf", "fp64", A), ("perm", "dd", P), ("tor", "certified", T), ("haf", "auto", np.zeros((0, 0)))]
    func, tier, dim, flat, off = pack(items)
    assert list(func) == [_lib.FUNCS[f] for f, _, _ in items]
    assert list(tier) == [_lib.TIERS["fp64"], _lib.TIERS["dd"], _lib.TIERS["certified"], _lib.GBSK_AUTO]
    assert list(dim) == [6, 4, 6, 0]                    # 5 -> 6: A (+) [1]
    assert list(off) == [0, 36, 52, 88]
    np.testing.assert_array_equal(flat[36:52].reshape(4, 4), P)
    aug = flat[:36].reshape(6, 6)
    np.testing.assert_array_equal(aug[:5, :5], A)
    assert aug[5, 5] == 1 and not aug[5, :5].any() and not aug[:5, 5].any()
    # the augmentation leaves the loop hafnian unchanged (the new vertex can only loop)
    assert oracle.lhaf_q(aug[None])[0] == pytest.approx(oracle.matchings(A, loop=True), rel=1e-12)


def test_pack_rejects_bad_items():
    from gbskernels_amd.serving import pack
    with pytest.raises(ValueError):
        pack([("tor", "fp64", np.eye(3))])              # odd torontonian
    with pytest.raises(ValueError):
        pack([("perm", "fp64", np.ones((2, 3)))])       # not square
    with pytest.raises(ValueError):
        pack([("det", "fp64", np.eye(2))])
    with pytest.raises(ValueError):
        pack([("perm", "quad", np.eye(2))])


def test_served_mix_stream():
    from gbskernels_amd.workloads import served_mix
    a, ta = served_mix(3000, seed=9)
    b, tb = served_mix(3000, seed=9)
    assert all(x[0] == y[0] and x[1] == y[1] and np.array_equal(x[2], y[2]) for x, y in zip(a, b))
    np.testing.assert_array_equal(ta, tb)
    funcs = {f for f, _, _ in a}
    assert funcs == {"perm", "haf", "lhaf"}
    dims = np.array([m.shape[0] for _, _, m in a])
    assert dims.min() >= 2 and dims.max() <= 24
    assert all(m.shape[0] % 2 == 0 for f, _, m in a if f != "perm")
    assert np.all(np.diff(ta) > 0)
    rate = len(ta) / ta[-1]
    assert 4.5e4 < rate < 5.5e4                          # Poisson 50k req/s
    # geometric(0.15) sizes: P(n = 2) = 0.15 / (1 - 0.85^23) for perm
    n_perm = dims[[f == "perm" for f, _, _ in a]]
    assert abs(np.mean(n_perm == 2) - 0.15 / (1 - 0.85 ** 23)) < 0.04


def test_canonical_flop_counts():
    """The canonical work counts the roofline divides by (perm 2^(n-1)(8n-4), the
    power-trace hafnian sum, the torontonian append count)."""
    import bench
    assert bench.perm_flops(16) == pytest.approx(4.06e6, rel=2e-3)
    assert bench.perm_flops(24) == pytest.approx(1.58e9, rel=2e-3)
    assert bench.perm_flops(28) == pytest.approx(2.95e10, rel=2e-3)
    assert bench.haf_flops(8) == pytest.approx(4.50e4, rel=1e-2)
    assert bench.haf_flops(16) == pytest.approx(1.02e7, rel=1e-2)
    assert bench.haf_flops(24) == pytest.approx(7.83e8, rel=1e-2)
    assert bench.haf_flops(24, loop=True) == pytest.approx(8.49e8, rel=1e-2)
    for nm, want in ((8, 3.87e4), (12, 1.31e6), (16, 3.61e7), (20, 8.88e8)):
        assert bench.tor_flops(nm) == pytest.approx(want, rel=2e-2)
    # 2^(n-1) (8n - 4) at n = 20
    assert math.isclose(bench.perm_flops(20), 2 ** 19 * 156)


def test_sieve_canonical_flops():
    """core/repeated.cu:86-132 per evaluated term: 4m^2 + 8m + 22K + 20."""
    import bench
    assert bench.sieve_flops((2,) * 6) == 365 * (4 * 36 + 48 + 22 * 6 + 20)
    assert bench.sieve_flops((6,) * 6) == 58825 * (4 * 36 + 48 + 22 * 18 + 20)
    assert bench.sieve_flops((0, 3, 0)) == 2 * (4 + 8 + 22 + 20)


def test_reference_table_inputs_and_fixture():
    """The reference-table leg's input recipes (core/bench_kernels.cu:58-72,
    306-326) and the published rows every compared cell maps to."""
    import json
    import bench
    from pathlib import Path
    m = bench.reference_inputs("sym", 8, 3)
    assert m.shape == (3, 8, 8) and np.array_equal(m, m.transpose(0, 2, 1))
    assert np.abs(m.real).max() <= 1.0 and np.abs(m.imag).max() <= 1.0 and np.abs(m.imag).max() > 0.5
    t = bench.reference_inputs("tor_real", 16, 2)
    assert np.array_equal(t, t.transpose(0, 2, 1)) and not t.imag.any() and np.abs(t).max() <= 0.05
    assert np.abs(bench.reference_inputs("tor", 8, 2)).max() <= 0.05 * np.sqrt(2)
    c = bench.reference_inputs("cert_sym", 8, 2)
    assert np.array_equal(c, c.transpose(0, 2, 1)) and np.abs(c.real).max() <= 0.3
    assert not np.array_equal(bench.reference_inputs("plain", 8, 1)[0], bench.reference_inputs("plain", 8, 1)[0].T)
    rows = json.loads((Path(bench.__file__).parent / "tests" / "golden" /
                       "kernel_throughput_reference.json").read_text())["rows"]
    assert all(r["evals_per_sec_median"] > 0 for r in rows)
    have = {(r["func"], r["matrix_dim"], r["gpu"]) for r in rows}
    for label, func, tier, kind, dims, ref_funcs in bench.REFERENCE_ROWS:
        for D in dims:
            for gpu in ("NVIDIA_GeForce_RTX_4090", "NVIDIA_A100-SXM4-40GB"):
                assert any((f, D, gpu) in have for f in ref_funcs), (label, D, gpu)
    # the published haf N = 8 medians (the reference's kernel table)
    got = {r["gpu"]: r["evals_per_sec_median"] for r in rows if r["func"] == "haf" and r["matrix_dim"] == 8}
    assert got == {"NVIDIA_GeForce_RTX_4090": 5562051.0, "NVIDIA_A100-SXM4-40GB": 10541870.0}


def test_executed_flop_table():
    """Counter-measured executed work: FP64 wave instructions x 64 lanes x at
    most 2 flops, and the helper's rates."""
    import bench
    for name, (flop, insts) in bench.EXECUTED_F64.items():
        assert insts * 64 <= flop <= insts * 128, name
    ex = bench.executed("perm_fp64_n24", 1024, 0.04)
    assert ex["tflops_executed"] == pytest.approx(2.014479e9 * 1024 / 0.04 / 1e12)
    assert 0.0 < ex["fp64_issue_frac"] < 1.0
    assert bench.executed("no_such_cell", 1, 1.0) == {}
