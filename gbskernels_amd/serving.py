"""Heterogeneous batches and the continuous-batching server (ctypes front of
``gbs_batch_*`` / ``gbs_server_*`` in ``include/gbsk.h``).

The reference evaluates a ragged request mix one uniform bucket at a time, each
bucket a synchronous host round trip (``gbskernels/__init__.py:813-852``).
:class:`Batch` evaluates a whole mix of functions, tiers and sizes in one call
(one packed copy in, buckets on a pool of streams, optionally replayed from a
captured hipGraph). :class:`Server` admits requests while earlier ones are on
the device and completes each bucket as soon as its kernels finish.

Items are ``(func, tier, matrix)`` with ``func`` in ``perm, haf, lhaf, tor``
and ``tier`` in ``fp64, dd, kappa, certified, dd_certified`` (plus ``auto`` for
the server). Matrices are square; odd loop hafnians are augmented with a unit
loop (``lhaf(A) == lhaf(A (+) [1])``) as the uniform API does. Values are
bit-identical to evaluating the same item alone through the uniform API.
"""

from __future__ import annotations

import ctypes
import os
import warnings
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib

SERVER_TIERS = {**_lib.TIERS, "auto": _lib.GBSK_AUTO}
_TIER_NAMES = {v: k for k, v in SERVER_TIERS.items()}


def _as_matrix(func: str, A) -> np.ndarray:
    M = np.ascontiguousarray(np.asarray(A, dtype=np.complex128))
    if M.ndim != 2 or M.shape[0] != M.shape[1]:
        raise ValueError(f"{func}: expected a square matrix, got shape {M.shape}")
    d = M.shape[0]
    if func == "lhaf" and d % 2:
        aug = np.zeros((d + 1, d + 1), np.complex128)
        aug[:d, :d] = M
        aug[d, d] = 1.0
        M = aug
    if func == "tor" and d % 2:
        raise ValueError("torontonian: matrix dimension must be even (2n x 2n)")
    return M


def pack(items: Sequence[tuple]):
    """Codes, dims, packed matrices and offsets of ``(func, tier, A)`` items."""
    n = len(items)
    func = np.empty(n, np.int32)
    tier = np.empty(n, np.int32)
    dim = np.empty(n, np.int32)
    off = np.empty(n, np.int64)
    mats = []
    pos = 0
    for i, (f, t, A) in enumerate(items):
        if f not in _lib.FUNCS:
            raise ValueError(f"unknown function {f!r}")
        if t not in SERVER_TIERS:
            raise ValueError(f"unknown tier {t!r}")
        M = _as_matrix(f, A)
        func[i], tier[i], dim[i], off[i] = _lib.FUNCS[f], SERVER_TIERS[t], M.shape[0], pos
        pos += M.size
        mats.append(M.ravel())
    flat = np.concatenate(mats) if mats else np.zeros(1, np.complex128)
    return func, tier, dim, np.ascontiguousarray(flat), off


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Batch:
    """Reusable handle for mixed batches (``gbs_batch_*``).

    ``budget_bytes`` caps the device bytes of one batch (0: no cap),
    ``nstreams`` is the stream pool, ``use_graphs`` captures the device part
    into a hipGraph cached by bucket signature.
    """

    def __init__(self, budget_bytes: int = 0, nstreams: int = 4, use_graphs: bool = True):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.gbs_batch_create(ctypes.byref(h), budget_bytes, nstreams, int(use_graphs)), "gbs_batch_create")
        self._h = h

    def run(self, items: Sequence[tuple]):
        """Evaluate ``(func, tier, A)`` items; returns ``(values, aux)``."""
        if self._h is None:
            raise RuntimeError("Batch is closed")
        if any(t == "auto" for _, t, _ in items):
            raise ValueError("tier 'auto' is served by Server; Batch runs kernel tiers")
        return self.run_packed(*pack(items))

    def run_packed(self, func, tier, dim, flat, off):
        """:meth:`run` on the arrays :func:`pack` returns."""
        if self._h is None:
            raise RuntimeError("Batch is closed")
        n = len(func)
        out = np.zeros(max(n, 1), np.complex128)
        aux = np.zeros(max(n, 1), np.float64)
        rc = _lib.lib().gbs_batch_run(self._h, n, _p(func), _p(tier), _p(dim), _p(flat), _p(off), _p(out), _p(aux))
        _lib.check(rc, "gbs_batch_run")
        return out[:n], aux[:n]

    def stats(self) -> dict:
        st = _lib.BatchStats()
        _lib.check(_lib.lib().gbs_batch_get_stats(self._h, ctypes.byref(st)), "gbs_batch_get_stats")
        return {k: int(getattr(st, k)) for k, _ in st._fields_}

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().gbs_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class RequestInfo:
    arrival_ns: int
    admitted_ns: int
    done_ns: int
    tier_used: str
    status: int

    @property
    def latency_s(self) -> float:
        return (self.done_ns - self.arrival_ns) * 1e-9


def _info(ri: _lib.RequestInfo) -> RequestInfo:
    return RequestInfo(int(ri.arrival_ns), int(ri.admitted_ns), int(ri.done_ns),
                       _TIER_NAMES.get(int(ri.tier_used), str(ri.tier_used)), int(ri.status))


class Server:
    """Continuous-batching server (``gbs_server_*``).

    A dispatcher thread admits everything pending into a free batch slot when
    its device bytes fit the budget left by the other slots (each slot's
    grow-only buffers count, busy or idle; idle slots give theirs back when a
    request would not fit otherwise); up to ``slots``
    batches overlap on the device; each bucket completes on its own event.
    ``auto`` requests run with kappa and are re-evaluated in double-double when
    ``sum|term| / |value| >= kappa_max``.

    The process should have at least ``slots * streams_per_slot`` HIP hardware
    queues (``GPU_MAX_HW_QUEUES``, read when HIP starts; its default is 4):
    streams sharing a queue run in order, so a light bucket waits behind a heavy
    kernel of another slot. A warning says so once per process.
    """

    _queue_warned = False

    def __init__(self, budget_bytes: int = 0, slots: int = 4, streams_per_slot: int = 4, max_batch: int = 0,
                 kappa_max: float = 1e8):
        try:
            hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        except ValueError:
            hwq = 4
        if hwq < slots * streams_per_slot and not Server._queue_warned:
            Server._queue_warned = True
            warnings.warn(f"GPU_MAX_HW_QUEUES={hwq} < {slots * streams_per_slot} server streams: streams share "
                          "hardware queues and light requests can wait behind heavy kernels; set it before the "
                          "first HIP call", RuntimeWarning, stacklevel=2)
        cfg = _lib.ServerConfig(budget_bytes, slots, streams_per_slot, max_batch, kappa_max)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().gbs_server_create(ctypes.byref(h), ctypes.byref(cfg)), "gbs_server_create")
        self._h = h

    @staticmethod
    def now_ns() -> int:
        return int(_lib.lib().gbs_now_ns())

    def submit(self, func: str, A, tier: str = "fp64", arrival_ns: int = 0) -> int:
        if self._h is None:
            raise RuntimeError("Server is closed")
        if func not in _lib.FUNCS or tier not in SERVER_TIERS:
            raise ValueError(f"unknown function/tier {func!r}/{tier!r}")
        M = _as_matrix(func, A)
        t = ctypes.c_longlong()
        rc = _lib.lib().gbs_server_submit(self._h, _lib.FUNCS[func], SERVER_TIERS[tier], M.shape[0], _p(M),
                                          int(arrival_ns), ctypes.byref(t))
        _lib.check(rc, "gbs_server_submit")
        return int(t.value)

    def wait(self, ticket: int):
        """``(value, aux, RequestInfo)`` of a submitted request (blocks)."""
        out = np.zeros(1, np.complex128)
        aux = np.zeros(1, np.float64)
        ri = _lib.RequestInfo()
        rc = _lib.lib().gbs_server_wait(self._h, ticket, _p(out), _p(aux), ctypes.byref(ri))
        _lib.check(rc, "gbs_server_wait")
        return complex(out[0]), float(aux[0]), _info(ri)

    def replay(self, items: Sequence[tuple], arrival_offsets_s: Sequence[float]):
        """Open-loop replay: item i is submitted at ``arrival_offsets_s[i]``
        seconds after the call starts. Returns ``(values, aux, infos)``."""
        return self.replay_packed(*pack(items), arrival_offsets_s)

    def replay_packed(self, func, tier, dim, flat, off, arrival_offsets_s):
        """:meth:`replay` on the arrays :func:`pack` returns."""
        if self._h is None:
            raise RuntimeError("Server is closed")
        n = len(func)
        arr = np.ascontiguousarray(np.round(np.asarray(arrival_offsets_s, np.float64) * 1e9).astype(np.int64))
        if arr.shape != (n,) or (n > 1 and np.any(np.diff(arr) < 0)):
            raise ValueError("arrival offsets must be ascending, one per item")
        out = np.zeros(max(n, 1), np.complex128)
        aux = np.zeros(max(n, 1), np.float64)
        infos = (_lib.RequestInfo * max(n, 1))()
        rc = _lib.lib().gbs_server_replay(self._h, n, _p(func), _p(tier), _p(dim), _p(flat), _p(off), _p(arr),
                                          _p(out), _p(aux), ctypes.cast(infos, ctypes.c_void_p))
        _lib.check(rc, "gbs_server_replay")
        return out[:n], aux[:n], [_info(infos[i]) for i in range(n)]

    def stats(self) -> dict:
        st = _lib.ServerStats()
        _lib.check(_lib.lib().gbs_server_get_stats(self._h, ctypes.byref(st)), "gbs_server_get_stats")
        return {k: int(getattr(st, k)) for k, _ in st._fields_}

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().gbs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def latency_percentiles(infos: Sequence[RequestInfo], qs=(50, 99)) -> dict:
    lat = np.array([i.latency_s for i in infos], np.float64)
    return {f"p{q}": float(np.percentile(lat, q)) for q in qs} if lat.size else {}
