"""ctypes binding of ``libgbsk.so`` (the C ABI declared in ``include/gbsk.h``).

There is no CPU fallback: if the library is missing or was built without the
HIP kernels, :func:`lib` raises. ctypes releases the GIL around each foreign
call, and the library serialises host calls on its own context, so Python
threads may call concurrently.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("GBSK_LIB", _HERE / "libgbsk.so"))

GBSK_ERR_ARG = 10001
GBSK_ERR_DIM = 10002
GBSK_ERR_FUNC = 10003
GBSK_ERR_WORKSPACE = 10004
GBSK_ERR_ALLOC = 10005
GBSK_ERR_SCRATCH = 10006

FUNCS = {"perm": 0, "haf": 1, "lhaf": 2, "tor": 3}
TIERS = {"fp64": 0, "dd": 1, "kappa": 2, "certified": 3, "dd_certified": 4}
GBSK_AUTO = 5  # server-only tier: kappa, re-evaluated in dd when ill-conditioned


class BatchStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_ulonglong)
                for k in ("batches", "items", "buckets", "graphs_captured", "graph_replays", "device_reallocs")]


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("slots", ctypes.c_int), ("streams_per_slot", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("kappa_max", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("arrival_ns", ctypes.c_longlong), ("admitted_ns", ctypes.c_longlong), ("done_ns", ctypes.c_longlong),
                ("tier_used", ctypes.c_int), ("status", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_ulonglong)
                for k in ("submitted", "completed", "batches", "buckets", "escalated", "max_in_flight",
                          "device_bytes_max")]

_cp = ctypes.c_void_p
_i = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)

# every symbol include/gbsk.h declares, with its ctypes signature
SIGNATURES = {
    **{f"gbs_{f}_host": ([_cp, _i, _i, _cp], _i) for f in ("perm", "perm_dd", "haf", "haf_dd", "lhaf", "lhaf_dd", "tor", "tor_dd")},
    **{f"gbs_{f}_{k}_host": ([_cp, _i, _i, _cp, _cp], _i)
       for f in ("perm", "haf", "lhaf", "tor") for k in ("kappa", "certified", "dd_certified")},
    "gbs_session_open": ([ctypes.POINTER(_cp)], _i),
    "gbs_session_evaluate": ([_cp, _i, _i, _cp, _i, _i, _cp], _i),
    "gbs_session_close": ([_cp], _i),
    "gbs_session_reallocs": ([_cp], ctypes.c_size_t),
    "gbs_session_evaluate_resident": ([_cp, _i, _i, _cp, _i, _i, ctypes.POINTER(_cp)], _i),
    "gbs_dev_to_host": ([_cp, _cp, _i], _i),
    "gbs_dev_free": ([_cp], _i),
    "gbs_workspace_bytes": ([_i, _i, _i, _i], ctypes.c_size_t),
    "gbs_eval_device": ([_i, _i, _cp, _i, _i, _cp, _cp, _cp, ctypes.c_size_t, _cp], _i),
    "gbs_eval_device_noalloc": ([_i, _i, _cp, _i, _i, _cp, _cp, _cp], _i),
    "gbs_max_dim": ([_i, _i], _i),
    "gbs_error_string": ([_i], ctypes.c_char_p),
    "gbs_is_device_build": ([], _i),
    "gbs_check_div_sqrt": ([_cp, _cp, ctypes.c_longlong, _cp, _cp, _cp], _i),
    "gbs_check_rsqrt": ([_cp, ctypes.c_longlong, _cp], _i),
    "gbs_batch_create": ([ctypes.POINTER(_cp), ctypes.c_size_t, _i, _i], _i),
    "gbs_batch_run": ([_cp, _i, _cp, _cp, _cp, _cp, _cp, _cp, _cp], _i),
    "gbs_batch_get_stats": ([_cp, ctypes.POINTER(BatchStats)], _i),
    "gbs_batch_destroy": ([_cp], _i),
    "gbs_now_ns": ([], ctypes.c_longlong),
    "gbs_server_create": ([ctypes.POINTER(_cp), ctypes.POINTER(ServerConfig)], _i),
    "gbs_server_submit": ([_cp, _i, _i, _i, _cp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)], _i),
    "gbs_server_wait": ([_cp, ctypes.c_longlong, _cp, _cp, _cp], _i),
    "gbs_server_replay": ([_cp, _i, _cp, _cp, _cp, _cp, _cp, _cp, _cp, _cp, _cp], _i),
    "gbs_server_get_stats": ([_cp, ctypes.POINTER(ServerStats)], _i),
    "gbs_server_destroy": ([_cp], _i),
    "gbs_lhaf_repeated_host": ([_cp, _cp, _i, _cp, _i, _cp], _i),
    "gbs_tor_recursive_single_host": ([_cp, _i, _i, _cp], _i),
    "gbs_tor_single_certified_host": ([_cp, _i, _i, _cp, _cp], _i),
    "gbs_tor_single_ddcertified_host": ([_cp, _i, _i, _cp, _cp], _i),
    "gbs_sampler_sample": ([_cp, _i, _cp, _i, _i, _i, _i, _cp, ctypes.c_ulonglong, _cp], _i),
    "gbs_philox4x32_10": ([_cp, _cp, _cp], None),
    "gbs_perm_get_lane_log2": ([_i, _i, _cp], _i),
    "gbs_perm_set_lane_log2": ([_i, _i, _i], _i),
    "gbs_perm_split": ([_i, _i, _cp, _cp], _i),
    "gbs_haf_get_schedule": ([_cp, _cp], _i),
    "gbs_haf_set_schedule": ([_i, _i], _i),
    "gbs_tor_get_schedule": ([_i, _cp, _cp], _i),
    "gbs_tor_set_schedule": ([_i, _i, _i], _i),
    "gbs_lhaf_repeated_cert_host": ([_cp, _cp, _i, _cp, _i, _cp, _cp], _i),
}

_lib = None


class GbskError(RuntimeError):
    """A nonzero return code from libgbsk."""

    def __init__(self, rc: int, where: str):
        self.rc = rc
        msg = lib().gbs_error_string(rc).decode()
        super().__init__(f"gbskernels: GPU error {rc} ({msg}) in {where}")


def _share_torch_hip_runtime() -> None:
    """PyTorch-ROCm wheels bundle their own HIP and HSA runtimes (same sonames
    as ROCm's). Two HSA runtimes in one process cannot both open the GPU: torch
    initialised after a ROCm-runtime library finds no devices. Loading torch's
    runtime first, by path and without importing torch, makes libgbsk's
    ``libamdhip64.so.7`` / ``libhsa-runtime64.so.1`` resolve to it, so device
    pointers, streams and DLPack hand-offs are shared with torch whichever is
    initialised first. ``GBSK_SYSTEM_HIP=1`` keeps the ROCm runtime instead."""
    if os.environ.get("GBSK_SYSTEM_HIP") == "1":
        return
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    for loc in (spec.submodule_search_locations or []) if spec else []:
        hip = Path(loc) / "lib" / "libamdhip64.so"
        if hip.exists():
            ctypes.CDLL(str(hip), mode=ctypes.RTLD_GLOBAL)
            return


def lib():
    """Load ``libgbsk.so`` once; raise if it is absent (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"gbskernels_amd: the HIP library {LIB_PATH} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    _share_torch_hip_runtime()
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.gbs_is_device_build() != 1:
        raise RuntimeError("gbskernels_amd: libgbsk.so is not a device build")
    _lib = L
    return L


def check(rc: int, where: str) -> None:
    if rc != 0:
        if rc in (GBSK_ERR_DIM, GBSK_ERR_ARG):
            raise ValueError(f"gbskernels: {lib().gbs_error_string(rc).decode()} in {where}")
        if rc == GBSK_ERR_FUNC:
            raise NotImplementedError(f"gbskernels: {lib().gbs_error_string(rc).decode()} in {where}")
        raise GbskError(rc, where)


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def max_dim(func: str, tier: str) -> int:
    return int(lib().gbs_max_dim(FUNCS[func], TIERS[tier]))
