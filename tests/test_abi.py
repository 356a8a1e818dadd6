"""CPU-side checks of the C ABI: libgbsk.so loads, exports every symbol that
include/gbsk.h declares, and its pure (no-GPU) queries behave. No compute call
is made here."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "gbsk.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|long long|const char\*)\s+(gbs_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_surface():
    syms = header_symbols()
    # the reference binding's extern "C" set (bindings/gbskernels_ext.cpp:24-68)
    for name in ["gbs_tor_recursive_single_host", "gbs_tor_single_certified_host",
                 "gbs_tor_single_ddcertified_host", "gbs_lhaf_repeated_host", "gbs_lhaf_repeated_cert_host",
                 "gbs_sampler_sample", "gbs_perm_host", "gbs_perm_dd_host", "gbs_haf_host", "gbs_haf_dd_host", "gbs_lhaf_host",
                 "gbs_lhaf_dd_host", "gbs_tor_host", "gbs_tor_dd_host", "gbs_perm_kappa_host",
                 "gbs_haf_kappa_host", "gbs_lhaf_kappa_host", "gbs_tor_kappa_host", "gbs_perm_certified_host",
                 "gbs_haf_certified_host", "gbs_lhaf_certified_host", "gbs_tor_certified_host",
                 "gbs_perm_dd_certified_host", "gbs_haf_dd_certified_host", "gbs_session_open",
                 "gbs_session_evaluate", "gbs_session_close", "gbs_session_reallocs",
                 "gbs_session_evaluate_resident", "gbs_dev_to_host", "gbs_dev_free"]:
        assert name in syms, name


def test_library_exports_every_declared_symbol():
    from gbskernels_amd import _lib
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding table covers the header exactly
    assert sorted(_lib.SIGNATURES) == header_symbols()


def test_pure_queries():
    from gbskernels_amd import _lib
    L = _lib.lib()
    assert L.gbs_is_device_build() == 1
    assert _lib.max_dim("perm", "fp64") == 28
    assert _lib.max_dim("perm", "dd") == 28
    assert _lib.max_dim("tor", "dd") == 40
    assert L.gbs_error_string(10002).decode().startswith("matrix dimension")
    # chunked permanents need workspace for partials, unchunked ones none
    assert L.gbs_workspace_bytes(0, 0, 16, 1000) == 0
    assert L.gbs_workspace_bytes(0, 0, 24, 1000) > 0


def test_perm_split_table_queries():
    """The permanent's chunk-shape table (host-side state, no GPU call)."""
    from gbskernels_amd import _lib
    L = _lib.lib()
    k, w, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for family in (0, 1):
        for n in (1, 8, 16, 24, 28):
            assert L.gbs_perm_split(family, n, ctypes.byref(k), ctypes.byref(w)) == 0
            # 2^(n-1) Gray terms = lanes per matrix x workgroups x terms per lane
            lanes = 1 << min(max(n - 1 - (2 if family else 4), 0), 6)
            assert lanes * w.value * (1 << k.value) == 1 << (n - 1)
    assert L.gbs_perm_get_lane_log2(0, 24, ctypes.byref(v)) == 0
    old = v.value
    try:
        assert L.gbs_perm_set_lane_log2(0, 24, 12) == 0
        assert L.gbs_perm_split(0, 24, ctypes.byref(k), ctypes.byref(w)) == 0
        assert (k.value, w.value) == (12, 32)
        # a larger chunk needs less workspace (fewer partials)
        assert L.gbs_workspace_bytes(0, 0, 24, 100) == 100 * 32 * 32
    finally:
        assert L.gbs_perm_set_lane_log2(0, 24, old) == 0
    ERR_ARG = 10001
    assert L.gbs_perm_set_lane_log2(0, 24, 3) == ERR_ARG      # below 2^4 terms per lane
    assert L.gbs_perm_set_lane_log2(1, 24, 1) == ERR_ARG
    assert L.gbs_perm_set_lane_log2(2, 24, 9) == ERR_ARG      # no such family
    assert L.gbs_perm_set_lane_log2(0, 29, 9) == ERR_ARG
    assert L.gbs_perm_split(0, 0, ctypes.byref(k), ctypes.byref(w)) == ERR_ARG


def test_ext_module_surface():
    import importlib
    import sys
    sys.path.insert(0, str(ROOT / "gbskernels_amd" / "ext"))
    ext = importlib.import_module("gbskernels_ext")
    assert ext.__host_shim__ is False
    for name in ["perm", "perm_dd", "haf", "haf_dd", "lhaf", "lhaf_dd", "tor", "tor_dd", "perm_kappa",
                 "haf_kappa", "lhaf_kappa", "tor_kappa", "perm_certified", "haf_certified", "lhaf_certified",
                 "tor_certified", "perm_dd_certified", "haf_dd_certified", "Session",
                 # optional attributes the reference probes with hasattr
                 "lhaf_dd_certified", "tor_dd_certified", "lhaf_repeated", "lhaf_repeated_certified",
                 "tor_single", "tor_single_certified", "tor_single_ddcertified", "sample_resident"]:
        assert hasattr(ext, name), name
    for m in ["perm", "haf", "lhaf", "tor", "perm_dd", "haf_dd", "lhaf_dd", "tor_dd", "perm_resident",
              "haf_resident", "lhaf_resident", "tor_resident", "close", "reallocs"]:
        assert hasattr(ext.Session, m), m
    with pytest.raises(ValueError, match="square"):
        ext.perm(__import__("numpy").zeros((2, 3, 4)))
    with pytest.raises(ValueError, match="even"):
        ext.tor(__import__("numpy").zeros((2, 3, 3)))


def test_api_argument_validation_without_gpu():
    import numpy as np
    import gbskernels_amd as g
    with pytest.raises(ValueError, match="certified"):
        g.perm_batched(np.zeros((1, 2, 2)), precision="certified")
    with pytest.raises(ValueError, match="rtol"):
        g.perm_batched(np.zeros((1, 2, 2)), precision="fp64", rtol=1e-3)
    with pytest.raises(NotImplementedError):
        g.perm_batched(np.zeros((1, 2, 2)), backend="cpu")
    with pytest.raises(ValueError, match="exceeds"):
        g.perm_batched(np.zeros((1, 29, 29)))
    with pytest.raises(ValueError, match="uniform"):
        g.perm_batched([np.zeros((2, 2)), np.zeros((3, 3))])
    # the mpmath ref tier is CPU-only, as in the reference
    assert g.perm(np.array([[1.0, 2.0], [3.0, 4.0]]), precision="ref", backend="cpu") == pytest.approx(10.0)
    assert g.lhaf(np.ones((3, 3)), precision="ref", backend="cpu") == pytest.approx(4.0)
    assert g.tor(0.2 * np.eye(4), precision="ref", backend="cpu") == pytest.approx(0.0625)
