"""ctypes front end of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, where it is the checker. The product
package ``gbskernels_amd`` never imports it.

Each function takes a numpy stack ``(B, d, d)`` (or one matrix) and returns
``(B,)`` complex128 (or float64 for sum|term| indicators and the real recursive
torontonian). The algorithms are restatements of the reference's ``cpu_ref``
(file:line citations live in ``gbsk_oracle.cpp``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "or_perm": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_perm_abs": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_perm_dd": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_perm_ryser": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_perm_q": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_powertrace": [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "or_matchings": [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    "or_tor": [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "or_tor_recursive": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_tor_dd": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_tor_q": [_dp, ctypes.c_int, ctypes.c_int, _dp],
    "or_lhaf_repeated": [_dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_int, _dp],
    "or_gen_bench_complex": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp],
    "or_gen_bench_real": [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp],
}


def build() -> None:
    """Compile ``liboracle.so`` with the committed Makefile (needs g++)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _stack(A, dtype=np.complex128) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(A, dtype=dtype))
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3 or a.shape[1] != a.shape[2]:
        raise ValueError(f"expected (B, d, d) or (d, d), got {a.shape}")
    return a


def set_threads(n: int) -> None:
    """OpenMP thread count for the batched oracle calls."""
    os.environ["OMP_NUM_THREADS"] = str(int(n))
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def perm(A) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_perm(_p(a), a.shape[1], a.shape[0], _p(out.view(np.float64)))
    return out


def perm_abs(A) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.float64)
    lib().or_perm_abs(_p(a), a.shape[1], a.shape[0], _p(out))
    return out


def perm_dd(A) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_perm_dd(_p(a), a.shape[1], a.shape[0], _p(out.view(np.float64)))
    return out


def perm_ryser(A) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_perm_ryser(_p(a), a.shape[1], a.shape[0], _p(out.view(np.float64)))
    return out


def perm_q(A) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_perm_q(_p(a), a.shape[1], a.shape[0], _p(out.view(np.float64)))
    return out


def _powertrace(A, loop: int, tier: int, want_abs: bool = False):
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    aux = np.empty(a.shape[0], np.float64) if want_abs else None
    lib().or_powertrace(_p(a), a.shape[1], a.shape[0], loop, tier,
                        _p(out.view(np.float64)), _p(aux) if aux is not None else None)
    return (out, aux) if want_abs else out


def haf(A):
    return _powertrace(A, 0, 0)


def haf_abs(A):
    return _powertrace(A, 0, 0, True)[1]


def haf_dd(A):
    return _powertrace(A, 0, 1)


def haf_q(A):
    return _powertrace(A, 0, 2)


def _even_lhaf(a: np.ndarray) -> np.ndarray:
    # odd N: lhaf(A) == lhaf(A (+) [1]) (a vertex that can only loop, weight 1)
    if a.shape[1] % 2 == 0:
        return a
    B, d, _ = a.shape
    aug = np.zeros((B, d + 1, d + 1), np.complex128)
    aug[:, :d, :d] = a
    aug[:, d, d] = 1.0
    return aug


def lhaf(A):
    return _powertrace(_even_lhaf(_stack(A)), 1, 0)


def lhaf_abs(A):
    return _powertrace(_even_lhaf(_stack(A)), 1, 0, True)[1]


def lhaf_dd(A):
    return _powertrace(_even_lhaf(_stack(A)), 1, 1)


def lhaf_q(A):
    return _powertrace(_even_lhaf(_stack(A)), 1, 2)


def matchings(A, loop: bool = False) -> np.ndarray:
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_matchings(_p(a), a.shape[1], a.shape[0], int(loop), _p(out.view(np.float64)))
    return out


def tor(A, want_abs: bool = False):
    a = _stack(A)
    if a.shape[1] % 2:
        raise ValueError("torontonian needs an even dimension")
    out = np.empty(a.shape[0], np.complex128)
    aux = np.empty(a.shape[0], np.float64) if want_abs else None
    lib().or_tor(_p(a), a.shape[1] // 2, a.shape[0], _p(out.view(np.float64)),
                 _p(aux) if aux is not None else None)
    return (out, aux) if want_abs else out


def tor_abs(A):
    return tor(A, True)[1]


def tor_recursive(A) -> np.ndarray:
    a = _stack(np.real(np.asarray(A)), np.float64)
    out = np.empty(a.shape[0], np.float64)
    lib().or_tor_recursive(_p(a), a.shape[1] // 2, a.shape[0], _p(out))
    return out


def tor_dd(A):
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_tor_dd(_p(a), a.shape[1] // 2, a.shape[0], _p(out.view(np.float64)))
    return out


def tor_q(A):
    a = _stack(A)
    out = np.empty(a.shape[0], np.complex128)
    lib().or_tor_q(_p(a), a.shape[1] // 2, a.shape[0], _p(out.view(np.float64)))
    return out


def gen_bench_complex(dim: int, batch: int, symmetric: bool, scale: float = 1.0) -> np.ndarray:
    """Inputs of the reference's kernel-bench cells (mt19937_64(1234 + dim))."""
    out = np.empty((batch, dim, dim), np.complex128)
    lib().or_gen_bench_complex(dim, batch, int(symmetric), float(scale), _p(out.view(np.float64)))
    return out


def gen_bench_real(dim: int, batch: int, scale: float) -> np.ndarray:
    out = np.empty((batch, dim, dim), np.float64)
    lib().or_gen_bench_real(dim, batch, float(scale), _p(out))
    return out


ABS = {"perm": perm_abs, "haf": haf_abs, "lhaf": lhaf_abs, "tor": tor_abs}
FP64 = {"perm": perm, "haf": haf, "lhaf": lhaf, "tor": tor}
DD = {"perm": perm_dd, "haf": haf_dd, "lhaf": lhaf_dd, "tor": tor_dd}
QUAD = {"perm": perm_q, "haf": haf_q, "lhaf": lhaf_q, "tor": tor_q}


def lhaf_repeated(A, gamma, reps, quad: bool = False) -> np.ndarray:
    """Repeated-row loop hafnian by the sieve (cpu_ref/repeated.py:61-122):
    base ``A`` (M, M), loop weights ``gamma`` (M,) (None -> zeros), counts
    ``reps`` (B, M) or (M,); returns (B,) complex128 (double or __float128
    arithmetic)."""
    a = np.ascontiguousarray(np.asarray(A, np.complex128))
    M = a.shape[0]
    g = np.zeros(M, np.complex128) if gamma is None else np.ascontiguousarray(np.asarray(gamma, np.complex128))
    r = np.ascontiguousarray(np.atleast_2d(np.asarray(reps, np.int32)))
    if a.shape != (M, M) or g.shape != (M,) or r.shape[1] != M or (r < 0).any():
        raise ValueError("lhaf_repeated: A (M, M), gamma (M,), reps (B, M) non-negative")
    out = np.empty(r.shape[0], np.complex128)
    lib().or_lhaf_repeated(_p(a.view(np.float64)), _p(g.view(np.float64)), M,
                           r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), r.shape[0], int(quad),
                           _p(out.view(np.float64)))
    return out


def expand_repeated(A, gamma, reps) -> np.ndarray:
    """The expanded matrix the sieve evaluates: row/column i of ``A`` repeated
    ``reps[i]`` times, diagonal gamma_i on every copy of i."""
    idx = np.repeat(np.arange(len(reps)), np.asarray(reps, np.int64))
    E = np.array(np.asarray(A, np.complex128)[np.ix_(idx, idx)])
    if len(idx):
        g = np.zeros(len(reps), np.complex128) if gamma is None else np.asarray(gamma, np.complex128)
        np.fill_diagonal(E, g[idx])
    return E
