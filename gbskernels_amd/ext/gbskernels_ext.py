"""``gbskernels_ext`` -- drop-in replacement of the reference's compiled module.

The reference's Python package imports a module named ``gbskernels_ext`` and
calls its attributes by name (reference ``gbskernels/__init__.py:109-136,224``).
This module provides the same attribute set (reference
``bindings/gbskernels_ext.cpp:231-391``) over the C ABI of ``libgbsk.so``:
put this directory on ``sys.path`` (or point ``GBSKERNELS_EXT_DIR`` at it) and
the reference package runs on the MI355X kernels unchanged.

Marshalling contract (reference ``bindings/gbskernels_ext.cpp:96-141``): input is
a ``(B, d, d)`` complex128 stack (coerced to C-contiguous complex128 here); the
kernel dimension is ``d`` for perm/haf/lhaf and ``d // 2`` for tor; a non-square
stack or an odd torontonian dimension raises ``ValueError``; a nonzero return
code raises ``RuntimeError``; results are fresh numpy arrays.
"""

from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_pkg_root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _pkg_root not in sys.path:
    sys.path.insert(0, _pkg_root)

from gbskernels_amd import _lib  # noqa: E402

__host_shim__ = False  # a real device build (reference: gpu_backend_kind)


def _stack(a, tor: bool) -> np.ndarray:
    s = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    if s.ndim != 3 or s.shape[1] != s.shape[2]:
        raise ValueError("gbskernels: each matrix must be square (B, d, d)")
    if tor and s.shape[1] % 2 != 0:
        raise ValueError("torontonian: matrix dimension must be even (2n x 2n)")
    return s


def _eval(name: str, a) -> np.ndarray:
    tor = name.startswith("tor")
    s = _stack(a, tor)
    B, d = s.shape[0], s.shape[1]
    out = np.empty(max(B, 1), np.complex128)[:B]
    fn = getattr(_lib.lib(), f"gbs_{name}_host")
    _lib.check(fn(_lib.ptr(s), d // 2 if tor else d, B, _lib.ptr(out)), name)
    return out


def _eval_aux(name: str, a):
    tor = name.startswith("tor")
    s = _stack(a, tor)
    B, d = s.shape[0], s.shape[1]
    out = np.empty(max(B, 1), np.complex128)[:B]
    aux = np.empty(max(B, 1), np.float64)[:B]
    fn = getattr(_lib.lib(), f"gbs_{name}_host")
    _lib.check(fn(_lib.ptr(s), d // 2 if tor else d, B, _lib.ptr(out), _lib.ptr(aux)), name)
    return out, aux


def perm(a): return _eval("perm", a)
def perm_dd(a): return _eval("perm_dd", a)
def haf(a): return _eval("haf", a)
def haf_dd(a): return _eval("haf_dd", a)
def lhaf(a): return _eval("lhaf", a)
def lhaf_dd(a): return _eval("lhaf_dd", a)
def tor(a): return _eval("tor", a)
def tor_dd(a): return _eval("tor_dd", a)

def perm_kappa(a): return _eval_aux("perm_kappa", a)
def haf_kappa(a): return _eval_aux("haf_kappa", a)
def lhaf_kappa(a): return _eval_aux("lhaf_kappa", a)
def tor_kappa(a): return _eval_aux("tor_kappa", a)

def perm_certified(a): return _eval_aux("perm_certified", a)
def haf_certified(a): return _eval_aux("haf_certified", a)
def lhaf_certified(a): return _eval_aux("lhaf_certified", a)
def tor_certified(a): return _eval_aux("tor_certified", a)
def perm_dd_certified(a): return _eval_aux("perm_dd_certified", a)
def haf_dd_certified(a): return _eval_aux("haf_dd_certified", a)
def lhaf_dd_certified(a): return _eval_aux("lhaf_dd_certified", a)
def tor_dd_certified(a): return _eval_aux("tor_dd_certified", a)


def _repeated_args(name: str, A, gamma, reps):
    """Shape checks of bindings/gbskernels_ext.cpp:296-347 (ValueError where the
    binding raises std::invalid_argument)."""
    a = np.ascontiguousarray(np.asarray(A, dtype=np.complex128))
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError(f"{name}: A must be (M, M)")
    M = a.shape[0]
    g = np.ascontiguousarray(np.asarray(gamma, dtype=np.complex128))
    if g.ndim != 1 or g.shape[0] != M:
        raise ValueError(f"{name}: gamma must be (M,)")
    r = np.ascontiguousarray(np.asarray(reps, dtype=np.int32))
    if r.ndim != 2 or r.shape[1] != M:
        raise ValueError(f"{name}: reps must be (B, M)")
    return a, g, r


def lhaf_repeated(A, gamma, reps):
    """Repeated-row loop hafnians: one (M, M) base matrix and (M,) loop vector,
    a (B, M) int32 repetition table -> (B,) values (finite-difference sieve)."""
    a, g, r = _repeated_args("lhaf_repeated", A, gamma, reps)
    B = r.shape[0]
    out = np.empty(max(B, 1), np.complex128)[:B]
    _lib.check(_lib.lib().gbs_lhaf_repeated_host(_lib.ptr(a), _lib.ptr(g), a.shape[0], _lib.ptr(r), B,
                                                 _lib.ptr(out)), "lhaf_repeated")
    return out


def lhaf_repeated_certified(A, gamma, reps):
    """Certified sieve: (values, bounds) with |value - exact| <= bound; values
    bit-identical to lhaf_repeated; over-cap rows give NaN and an inf bound."""
    a, g, r = _repeated_args("lhaf_repeated_certified", A, gamma, reps)
    B = r.shape[0]
    out = np.empty(max(B, 1), np.complex128)[:B]
    bnd = np.empty(max(B, 1), np.float64)[:B]
    _lib.check(_lib.lib().gbs_lhaf_repeated_cert_host(_lib.ptr(a), _lib.ptr(g), a.shape[0], _lib.ptr(r), B,
                                                      _lib.ptr(out), _lib.ptr(bnd)), "lhaf_repeated_certified")
    return out, bnd


def _tor_single_arg(name: str, O) -> np.ndarray:
    """Checks of bindings/gbskernels_ext.cpp:257-295: real (2n, 2n), 2n <= 64."""
    M = np.ascontiguousarray(np.asarray(O, dtype=np.float64))
    if M.ndim != 2 or M.shape[0] != M.shape[1] or M.shape[0] % 2 != 0 or M.shape[0] > 64:
        raise ValueError(f"{name}: O must be (2n, 2n) real, 2n <= 64")
    return M


def tor_single(O, g):
    """Single-large torontonian of one real (2n, 2n) O (2n <= 64) across the
    whole GPU; NaN off the physical (SPD) domain."""
    M = _tor_single_arg("tor_single", O)
    out = ctypes.c_double(0.0)
    _lib.check(_lib.lib().gbs_tor_recursive_single_host(_lib.ptr(M), M.shape[0] // 2, int(g),
                                                        ctypes.addressof(out)), "tor_single")
    return out.value


def _tor_single_cert(name: str, fn: str, O, g):
    M = _tor_single_arg(name, O)
    out, bnd = ctypes.c_double(0.0), ctypes.c_double(0.0)
    _lib.check(getattr(_lib.lib(), fn)(_lib.ptr(M), M.shape[0] // 2, int(g), ctypes.addressof(out),
                                       ctypes.addressof(bnd)), name)
    return out.value, bnd.value


def tor_single_certified(O, g):
    """Certified single-large torontonian -> (value, bound); NaN / inf off the domain."""
    return _tor_single_cert("tor_single_certified", "gbs_tor_single_certified_host", O, g)


def tor_single_ddcertified(O, g):
    """Certified double-double single-large torontonian -> (value, bound)."""
    return _tor_single_cert("tor_single_ddcertified", "gbs_tor_single_ddcertified_host", O, g)


def sample_resident(ak, off, M, num_draws, cutoff, maxn, invfac, seed):
    """Resident chain-rule GBS sampler (bindings/gbskernels_ext.cpp:153-167):
    ``ak`` the M reduced A-matrices concatenated (complex128), ``off`` their
    int32 offsets, ``invfac[j] = 1/j!``; returns (num_draws, M) int32 samples."""
    a = np.ascontiguousarray(np.asarray(ak, dtype=np.complex128).reshape(-1))
    o = np.ascontiguousarray(np.asarray(off, dtype=np.int32).reshape(-1))
    f = np.ascontiguousarray(np.asarray(invfac, dtype=np.float64).reshape(-1))
    M, D, C = int(M), int(num_draws), int(cutoff)
    if o.shape[0] < M or f.shape[0] < C + 1:
        raise ValueError("sample_resident: off needs M entries and invfac cutoff + 1")
    out = np.zeros((D, M), np.int32)
    _lib.check(_lib.lib().gbs_sampler_sample(_lib.ptr(a), a.shape[0], _lib.ptr(o), M, D, C, int(maxn), _lib.ptr(f),
                                             int(seed) & 0xFFFFFFFFFFFFFFFF, _lib.ptr(out)), "sample_resident")
    return out


# ---- DLPack export of device results (the reference returns an nb::ndarray
# that exposes __dlpack__, bindings/gbskernels_ext.cpp:209-228) ----------------
class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DLDeleter = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", _DLDeleter)]

KDL_ROCM = 10      # DLPack device type of HIP device memory
_KDL_COMPLEX = 5
_exported = {}     # managed-tensor address -> (struct, shape array, device pointer)


def _dl_release(addr: int) -> None:
    entry = _exported.pop(addr, None)
    if entry is not None and entry[2]:
        _lib.lib().gbs_dev_free(entry[2])


@_DLDeleter
def _dl_deleter(mt):
    _dl_release(ctypes.addressof(mt.contents))


_api = ctypes.pythonapi
_api.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_api.PyCapsule_New.restype = ctypes.py_object
_api.PyCapsule_IsValid.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_api.PyCapsule_IsValid.restype = ctypes.c_int
_api.PyCapsule_GetPointer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_api.PyCapsule_GetPointer.restype = ctypes.c_void_p


@ctypes.CFUNCTYPE(None, ctypes.c_void_p)
def _capsule_destructor(cap):
    # a capsule nobody consumed keeps the name "dltensor": free its tensor here
    # (a consumer renames it "used_dltensor" and calls the deleter itself)
    if _api.PyCapsule_IsValid(cap, b"dltensor"):
        _dl_release(_api.PyCapsule_GetPointer(cap, b"dltensor"))


class DeviceResult:
    """A (B,) complex128 result left in device memory (reference
    ``bindings/gbskernels_ext.cpp:209-228``). Owns the buffer; ``numpy()`` copies
    it to the host; ``ptr`` is the device address. ``__dlpack__`` hands the
    buffer zero-copy to a DLPack consumer (``torch.from_dlpack(r)`` gives a
    complex128 tensor on the GPU) and moves ownership to it; the values are
    complete when the result is returned, so any consumer stream may read them."""

    def __init__(self, ptr: int, n: int):
        self.ptr = ptr
        self.shape = (n,)
        self.dtype = np.dtype(np.complex128)

    def __len__(self) -> int:
        return self.shape[0]

    def numpy(self) -> np.ndarray:
        out = np.empty(max(self.shape[0], 1), np.complex128)[: self.shape[0]]
        if self.shape[0]:
            _lib.check(_lib.lib().gbs_dev_to_host(self.ptr, _lib.ptr(out), self.shape[0]), "dev_to_host")
        return out

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    def __dlpack_device__(self):
        return (KDL_ROCM, 0)

    def __dlpack__(self, stream=None, max_version=None, dl_device=None, copy=None):
        if self.ptr is None and self.shape[0]:
            raise BufferError("DeviceResult: the buffer was already exported or freed")
        mt = _DLManagedTensor()
        shape = (ctypes.c_int64 * 1)(self.shape[0])
        mt.dl_tensor.data = self.ptr
        mt.dl_tensor.device = _DLDevice(KDL_ROCM, 0)
        mt.dl_tensor.ndim = 1
        mt.dl_tensor.dtype = _DLDataType(_KDL_COMPLEX, 128, 1)
        mt.dl_tensor.shape = shape
        mt.dl_tensor.strides = None
        mt.dl_tensor.byte_offset = 0
        mt.manager_ctx = None
        mt.deleter = _dl_deleter
        _exported[ctypes.addressof(mt)] = (mt, shape, self.ptr)
        self.ptr = None  # the capsule (then its consumer) owns the buffer now
        return _api.PyCapsule_New(ctypes.addressof(mt), b"dltensor", ctypes.cast(_capsule_destructor, ctypes.c_void_p))

    def __del__(self):
        p, self.ptr = getattr(self, "ptr", None), None
        if p:
            try:
                _lib.lib().gbs_dev_free(p)
            except Exception:
                pass


_SESSION_FUNCS = {"perm": 0, "haf": 1, "lhaf": 2, "tor": 3}


class Session:
    """Device-resident session: grow-only buffers and one HIP stream reused by
    every call (reference ``core/host_api.cu:474-606``)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().gbs_session_open(ctypes.byref(h)), "session_open")
        self._h = h

    def _call(self, name: str, a):
        if self._h is None:
            raise RuntimeError("gbskernels: session is closed")
        func, _, prec = name.partition("_")
        s = _stack(a, func == "tor")
        B, d = s.shape[0], s.shape[1]
        out = np.empty(max(B, 1), np.complex128)[:B]
        _lib.check(_lib.lib().gbs_session_evaluate(self._h, _SESSION_FUNCS[func], 1 if prec == "dd" else 0,
                                                   _lib.ptr(s), d, B, _lib.ptr(out)), f"Session.{name}")
        return out

    def _resident(self, name: str, a) -> DeviceResult:
        if self._h is None:
            raise RuntimeError("gbskernels: session is closed")
        func, _, prec = name.partition("_")
        s = _stack(a, func == "tor")
        B, d = s.shape[0], s.shape[1]
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().gbs_session_evaluate_resident(self._h, _SESSION_FUNCS[func], 1 if prec == "dd" else 0,
                                                            _lib.ptr(s), d, B, ctypes.byref(p)), f"Session.{name}_resident")
        return DeviceResult(p.value or 0, B)

    def perm(self, a): return self._call("perm", a)
    def perm_dd(self, a): return self._call("perm_dd", a)
    def haf(self, a): return self._call("haf", a)
    def haf_dd(self, a): return self._call("haf_dd", a)
    def lhaf(self, a): return self._call("lhaf", a)
    def lhaf_dd(self, a): return self._call("lhaf_dd", a)
    def tor(self, a): return self._call("tor", a)
    def tor_dd(self, a): return self._call("tor_dd", a)
    def perm_resident(self, a): return self._resident("perm", a)
    def haf_resident(self, a): return self._resident("haf", a)
    def lhaf_resident(self, a): return self._resident("lhaf", a)
    def tor_resident(self, a): return self._resident("tor", a)

    def reallocs(self) -> int:
        return int(_lib.lib().gbs_session_reallocs(self._h)) if self._h is not None else 0

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().gbs_session_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
