"""gbskernels_amd -- MI355X (gfx950) batched backend for the permanent, hafnian,
loop hafnian and torontonian, behind the reference GBSKernels Python API.

Public surface (mirrors the reference ``gbskernels/__init__.py:54-63,517-558,734-908``
for ``backend="gpu"``): ``perm/haf/lhaf/tor`` and ``*_batched`` with
``precision in {"fp64", "dd", "auto", "certified", "ref"}``, ``return_diagnostics``
and ``rtol``; :class:`Workspace` for ragged, device-resident batches;
:func:`gpu_available` and :func:`gpu_backend_kind`; :func:`lhaf_repeated`, the
repeated-row loop hafnian of the sampling workload (reference
``gbskernels/__init__.py:561-632``).

Tiers (semantics of the reference, reference ``gbskernels/__init__.py:1-37``):

* ``fp64`` -- one launch of the hand-written HIP kernel per uniform stack;
* ``dd`` -- the same walk in double-double;
* ``auto`` -- fp64 plus the on-device cancellation indicator
  ``kappa = sum|term| / |value|``; elements with ``kappa >= 1e8`` are re-run in
  ``dd`` on the GPU (``ref`` beyond the dd domain);
* ``certified`` -- the fp64 value (bit-identical to ``fp64``) plus a rigorous
  bound; ``rtol`` escalates per element through certified-dd, then ``ref``;
* ``ref`` -- mpmath ground truth (CPU, tiny sizes), as in the reference.

``backend="cpu"`` serves only ``precision="ref"``: the reference's numpy CPU
tiers are not part of this package. Nothing here falls back to the CPU
silently: a missing HIP library raises.
"""

from __future__ import annotations

from typing import Any, Iterable

import numpy as np

from . import _lib, _ref

__version__ = "0.1.0"
__all__ = [
    "perm", "perm_batched", "haf", "haf_batched", "lhaf", "lhaf_batched",
    "tor", "tor_batched", "gpu_available", "gpu_backend_kind", "Workspace", "PRECISIONS",
    "lhaf_repeated", "sieve_term_count", "tor_single",
]

PRECISIONS = ("fp64", "ref", "dd", "auto", "certified")
_FUNCS = ("perm", "haf", "lhaf", "tor")
# reference gbskernels/__init__.py:240 (calibrated there; kept as the same contract)
_AUTO_KAPPA_MAX = 1e8


def _ext():
    from .ext import gbskernels_ext as ext
    _lib.lib()  # raise now if the HIP library is missing
    return ext


def gpu_available() -> bool:
    """True when ``libgbsk.so`` loads (it always carries device code)."""
    try:
        _lib.lib()
        return True
    except (OSError, RuntimeError):
        return False


def gpu_backend_kind() -> str:
    """``"gpu"`` for the loaded HIP library, ``"none"`` if it is absent
    (reference ``gbskernels/__init__.py:699-713``; there is no host-shim build)."""
    return "gpu" if gpu_available() else "none"


def _max_dim(func: str, tier: str) -> int:
    return _lib.max_dim(func, tier)


# --- validation (reference gbskernels/__init__.py:158-219) ---------------------

def _prepare(func: str, matrices: Iterable[Any], tier: str) -> np.ndarray:
    """Uniform (B, d, d) complex128 stack for one launch; augments odd loop
    hafnians; enforces the kernel's size range."""
    try:
        stack = np.ascontiguousarray(np.asarray(matrices, dtype=np.complex128))
    except ValueError:
        stack = None
    if stack is None or stack.ndim != 3 or stack.shape[1] != stack.shape[2]:
        raise ValueError(
            "GPU backend needs a uniform (B, d, d) stack of square matrices; "
            "ragged batches go through gbskernels_amd.Workspace.")
    d = int(stack.shape[1])
    if func == "lhaf" and d % 2:
        # lhaf(A) == lhaf(A (+) [1]): the added vertex can only loop, weight 1
        aug = np.zeros((stack.shape[0], d + 1, d + 1), np.complex128)
        aug[:, :d, :d] = stack
        aug[:, d, d] = 1.0
        stack, d = np.ascontiguousarray(aug), d + 1
    if func == "tor" and d % 2:
        raise ValueError("torontonian: matrix dimension must be even (2n x 2n)")
    cap = _max_dim(func, tier)
    if d > cap:
        raise ValueError(
            f"GPU {func} ({tier}): matrix dimension {d} exceeds the kernel's limit {cap}"
            + (" (an odd loop hafnian is augmented to N+1)" if func == "lhaf" else ""))
    if func == "tor" and tier in ("dd", "dd_certified") and stack.size \
            and float(np.max(np.abs(stack.imag))) > 1e-12:
        raise ValueError("GPU torontonian (precision='dd') is real-domain only; the input "
                         "has non-negligible imaginary parts")
    return stack


def _run(func: str, tier: str, stack: np.ndarray):
    ext = _ext()
    name = func if tier == "fp64" else f"{func}_{tier}"
    return getattr(ext, name)(stack)


# --- auto (reference gbskernels/__init__.py:243-297) ---------------------------

def _rerun_risky(func: str, stack: np.ndarray, values: np.ndarray, risky: np.ndarray) -> str:
    sub = np.ascontiguousarray(stack[risky])
    try:
        sub = _prepare(func, sub, "dd")
    except ValueError:
        for i in np.nonzero(risky)[0]:
            values[i] = _ref.evaluate(func, stack[i])
        return "ref"
    values[risky] = _run(func, "dd", sub)
    return "dd"


def _auto_batched(func: str, matrices, return_diagnostics: bool):
    stack = _prepare(func, matrices, "kappa")
    values, absnorm = _run(func, "kappa", stack)
    values = values.astype(np.complex128, copy=True)
    kappa = absnorm / np.maximum(np.abs(values), 1e-300)
    risky = kappa >= _AUTO_KAPPA_MAX
    tier = _rerun_risky(func, stack, values, risky) if risky.any() else None
    if not return_diagnostics:
        return values
    diags = [{"tier": tier if r else "fp64", "cancellation": float(k)} for r, k in zip(risky, kappa)]
    return values, diags


# --- certified (contract of reference gbskernels/__init__.py:350-478) ----------
#
# The whole batch is judged at once: one certified-fp64 launch, a vectorised
# test of every certificate against rtol, ONE certified-dd launch for all the
# elements the fp64 bounds cannot prove, and the mpmath ref tier only for what
# that still leaves. Each step is triggered by a rigorous bound, never by a
# heuristic. Diagnostic keys and messages are the reference's.

def _finite_input_or_raise(A) -> None:
    try:
        ok = bool(np.isfinite(np.asarray(A)).all())
    except (TypeError, ValueError):
        ok = False
    if not ok:
        raise ValueError("certified evaluation requires finite input entries")


class _Certificates:
    """Values and bounds of one certified launch with, per element, whether the
    certificate is usable (finite value, finite non-negative bound) and its
    relative bound (+inf when unusable; 0 / 0 counts as 0)."""

    def __init__(self, values, bounds) -> None:
        self.values = np.array(values, dtype=np.complex128)
        self.bounds = np.array(bounds, dtype=np.float64)
        v, b = self.values, self.bounds
        self.usable = np.isfinite(v.real) & np.isfinite(v.imag) & np.isfinite(b) & (b >= 0.0)
        mag = np.abs(v)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(mag > 0.0, b / np.where(mag > 0.0, mag, 1.0), np.where(b == 0.0, 0.0, np.inf))
        self.rel = np.where(self.usable, rel, np.inf)

    def proves(self, rtol: float) -> np.ndarray:
        return self.usable & (self.rel <= rtol)

    def diag(self, i: int, tier: str) -> dict:
        return {"tier": tier, "abs_error_bound": float(self.bounds[i]), "rel_error_bound": float(self.rel[i])}


def _dd_certificates(func: str, stack: np.ndarray):
    """Certified-dd certificates of a uniform stack, or None where that tier
    cannot take it (size beyond its kernels, complex torontonian input)."""
    try:
        sub = _prepare(func, stack, "dd_certified")
    except (ValueError, NotImplementedError):
        return None
    return _Certificates(*_run(func, "dd_certified", sub))


def _certified_batched(func: str, matrices, rtol):
    stack = _prepare(func, matrices, "certified")
    _finite_input_or_raise(stack)
    fp = _Certificates(*_run(func, "certified", stack))
    if rtol is None:
        refused = np.flatnonzero(~fp.usable)
        if refused.size:
            raise FloatingPointError(
                f"certified {func} batch element {int(refused[0])} refused with a non-finite value or bound")
        return fp.values, [fp.diag(i, "certified-fp64") for i in range(len(fp.values))]
    out = fp.values.copy()
    diags = [dict(fp.diag(i, "certified-fp64"), escalated=False) for i in range(len(out))]
    todo = np.flatnonzero(~fp.proves(rtol))
    if todo.size:
        dd = _dd_certificates(func, stack[todo])
        settled = dd.proves(rtol) if dd is not None else np.zeros(todo.size, bool)
        for k, i in enumerate(todo):
            if settled[k]:
                out[i] = dd.values[k]
                diags[i] = dict(dd.diag(k, "certified-dd"), escalated=True)
                continue
            ref = _ref.evaluate(func, stack[i])
            if not (np.isfinite(ref.real) and np.isfinite(ref.imag)):
                raise FloatingPointError(f"reference {func} evaluator returned a non-finite value")
            out[i] = ref
            # the diagnostics keep the fp64 certificate that forced the escalation
            diags[i] = {"tier": "ref", "escalated": True,
                        "abs_error_bound": float(fp.bounds[i]) if fp.usable[i] else float("inf"),
                        "rel_error_bound": float(fp.rel[i])}
    return out, diags


# --- dispatch (reference gbskernels/__init__.py:481-514) -----------------------

def _check_args(precision, backend, return_diagnostics, rtol):
    if precision not in PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; expected one of {PRECISIONS}")
    if backend not in ("cpu", "gpu"):
        raise ValueError(f"unknown backend {backend!r}; expected 'cpu' or 'gpu'")
    if return_diagnostics and precision not in ("auto", "certified"):
        raise ValueError("return_diagnostics=True requires precision='auto' or 'certified'")
    if precision == "certified" and not return_diagnostics:
        raise ValueError("precision='certified' returns its error bound via diagnostics; call with "
                         "return_diagnostics=True (a certificate you cannot see is not a certificate)")
    if rtol is not None and precision != "certified":
        raise ValueError("rtol is a certified-tier parameter; it requires precision='certified'")
    if rtol is not None and (not np.isfinite(rtol) or rtol < 0.0):
        raise ValueError("rtol must be finite and non-negative")
    if precision == "ref" and backend == "gpu":
        raise NotImplementedError("'ref' is the mpmath ground truth -- CPU only (backend='cpu')")
    if backend == "cpu" and precision != "ref":
        raise NotImplementedError(
            "gbskernels_amd is the GPU backend: backend='cpu' serves precision='ref' only "
            "(the numpy CPU tiers are the reference package's cpu_ref)")


def _dispatch_batched(func, matrices, precision, backend, return_diagnostics=False, rtol=None):
    _check_args(precision, backend, return_diagnostics, rtol)
    if precision == "ref":
        batch = matrices if isinstance(matrices, np.ndarray) and matrices.ndim == 3 else list(matrices)
        return np.array([_ref.evaluate(func, A) for A in batch], dtype=np.complex128)
    if precision == "auto":
        return _auto_batched(func, matrices, return_diagnostics)
    if precision == "certified":
        return _certified_batched(func, matrices, rtol)
    return np.asarray(_run(func, precision, _prepare(func, matrices, precision)))


def _dispatch_single(func, A, precision, backend, return_diagnostics=False, rtol=None):
    _check_args(precision, backend, return_diagnostics, rtol)
    if precision == "ref":
        return _ref.evaluate(func, A)
    if precision == "certified":
        _finite_input_or_raise(A)
    res = _dispatch_batched(func, np.asarray(A)[None, ...], precision, backend, return_diagnostics, rtol)
    if return_diagnostics:
        vals, diags = res
        return complex(vals[0]), diags[0]
    return complex(res[0])


def perm(A, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Permanent of one square matrix."""
    return _dispatch_single("perm", A, precision, backend, return_diagnostics, rtol)


def perm_batched(matrices, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Permanents of a uniform (B, n, n) stack -> complex128 (B,)."""
    return _dispatch_batched("perm", matrices, precision, backend, return_diagnostics, rtol)


def haf(A, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Hafnian of one symmetric matrix (the diagonal is ignored)."""
    return _dispatch_single("haf", A, precision, backend, return_diagnostics, rtol)


def haf_batched(matrices, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Hafnians of a uniform (B, N, N) stack of symmetric matrices."""
    return _dispatch_batched("haf", matrices, precision, backend, return_diagnostics, rtol)


def lhaf(A, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Loop hafnian of one symmetric matrix (diagonal = loop weights)."""
    return _dispatch_single("lhaf", A, precision, backend, return_diagnostics, rtol)


def lhaf_batched(matrices, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Loop hafnians of a uniform stack; odd N is augmented to N + 1."""
    return _dispatch_batched("lhaf", matrices, precision, backend, return_diagnostics, rtol)


def tor(O, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Torontonian of one 2n x 2n matrix O (xxpp ordering)."""
    return _dispatch_single("tor", O, precision, backend, return_diagnostics, rtol)


def tor_batched(matrices, precision: str = "fp64", backend: str = "gpu", return_diagnostics: bool = False, rtol=None):
    """Torontonians of a uniform (B, 2n, 2n) stack."""
    return _dispatch_batched("tor", matrices, precision, backend, return_diagnostics, rtol)


# --- repeated-row loop hafnian (reference gbskernels/__init__.py:561-632) --------

def sieve_term_count(reps: Any) -> int:
    """``prod(n_i + 1)``: the sieve's term count for one pattern (reference
    ``cpu_ref/repeated.py:54-58``)."""
    n = np.asarray(reps, dtype=np.int64)
    return int(np.prod(n + 1)) if n.size else 1


def lhaf_repeated(A: Any, gamma: Any, reps: Any, backend: str = "gpu", certified: bool = False):
    """Loop hafnians of ``A`` expanded by repetition patterns.

    ``A`` is the (M, M) base matrix, ``gamma`` the length-M loop weights
    (``None``: zeros, the ordinary hafnian of the expansion) and ``reps`` a
    (B, M) table of non-negative counts, or one length-M pattern. Row/column
    ``i`` is repeated ``reps[b, i]`` times with ``gamma[i]`` on each copy's
    diagonal; the finite-difference sieve costs ``prod(n_i + 1)`` terms per
    pattern. Returns (B,) values (a scalar for one pattern); with
    ``certified=True`` also ``{"tier": "certified-fp64", "abs_error_bound": ...}``
    with ``|value - exact| <= bound`` and values bit-identical to the plain
    sieve. Patterns beyond the kernel's domain (M > 32, N > 64, more than 2^40
    terms) come back NaN; certified ones then raise ``FloatingPointError``."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"lhaf_repeated: A must be square (M, M), got {A.shape}")
    M = A.shape[0]
    g = np.zeros(M, dtype=np.complex128) if gamma is None else np.ascontiguousarray(gamma, dtype=np.complex128)
    if g.shape != (M,):
        raise ValueError(f"lhaf_repeated: gamma must be a length-{M} vector, got {g.shape}")
    if certified:
        _finite_input_or_raise(A)
        _finite_input_or_raise(g)
    r = np.ascontiguousarray(reps, dtype=np.int32)
    single = r.ndim == 1
    if single:
        r = r[None, :]
    if r.ndim != 2 or r.shape[1] != M:
        raise ValueError(f"lhaf_repeated: reps must be (B, {M}) or ({M},), got {np.asarray(reps).shape}")
    if (r < 0).any():
        raise ValueError("lhaf_repeated: reps must be non-negative")
    if backend not in ("cpu", "gpu"):
        raise ValueError(f"unknown backend {backend!r}; expected 'cpu' or 'gpu'")
    if backend == "cpu":
        raise NotImplementedError(
            "gbskernels_amd is the GPU backend: the CPU sieve is the reference's cpu_ref.repeated")
    ext = _ext()
    if certified:
        v, bnd = ext.lhaf_repeated_certified(A, g, r)
        valid = bool(np.all(np.isfinite(v.real)) and np.all(np.isfinite(v.imag))
                     and np.all(np.isfinite(bnd)) and np.all(bnd >= 0.0))
        if not valid:
            raise FloatingPointError("certified lhaf_repeated evaluator refused with a non-finite value or invalid bound")
        diag = {"tier": "certified-fp64", "abs_error_bound": float(bnd[0]) if single else bnd}
        return (complex(v[0]) if single else v), diag
    v = ext.lhaf_repeated(A, g, r)
    return complex(v[0]) if single else v


# --- single-large torontonian (contract of reference gbskernels/__init__.py:635-691)

# (certified, dd) -> extension entry point and the tier label of its certificate
_TOR_SINGLE_ENTRY = {
    (False, False): ("tor_single", None),
    (True, False): ("tor_single_certified", "certified-fp64"),
    (True, True): ("tor_single_ddcertified", "certified-dd"),
    (False, True): ("tor_single_ddcertified", "certified-dd"),
}


def _single_tor_matrix(O: Any) -> np.ndarray:
    """The binary64 matrix the single-large kernels walk: real (a complex array
    must have zero imaginary parts), square, even, dim <= 64, finite and exactly
    symmetric (the recursive Cholesky reads one triangle)."""
    raw = np.asarray(O)
    if np.iscomplexobj(raw) and np.any(raw.imag != 0):
        raise ValueError("tor_single is real-domain only; use tor() for complex O")
    M = np.ascontiguousarray(raw.real if np.iscomplexobj(raw) else raw, dtype=np.float64)
    d = M.shape[0] if M.ndim == 2 else -1
    problems = (
        (M.ndim != 2 or M.shape[0] != M.shape[1] or d % 2 != 0,
         f"tor_single: O must be square with even dimension, got {M.shape}"),
        (d > 64, f"tor_single: dimension {d} exceeds the cap 64 (n <= 32 modes)"),
        (not bool(np.isfinite(M).all()), "tor_single: O must contain only finite binary64 values"),
        (not np.array_equal(M, M.T), "tor_single: O must be exactly symmetric after conversion to binary64"),
    )
    for failed, message in problems:
        if failed:
            raise ValueError(message)
    return M


def tor_single(O: Any, groups: int | None = None, certified: bool = False, dd: bool = False):
    """Torontonian of ONE real ``(2n, 2n)`` matrix, ``2n <= 64``, split across
    the whole GPU. Physical-domain inputs (``I - O_S`` SPD) only: off-domain
    raises ``ValueError``. ``groups`` (default ``min(n, 14)``) is validated for
    compatibility; the work split is this backend's own and the value does not
    depend on it. ``certified`` also returns ``{"tier", "abs_error_bound",
    "rel_error_bound"}``; ``dd=True`` (implies certified) carries the value in
    double-double."""
    M = _single_tor_matrix(O)
    n = M.shape[0] // 2
    g = min(n, 14) if groups is None else int(groups)
    if g < 0 or g > n:
        raise ValueError(f"groups must be in [0, {n}]")
    entry, tier = _TOR_SINGLE_ENTRY[(bool(certified), bool(dd))]
    res = getattr(_ext(), entry)(M, g)
    if tier is None:
        value = float(res)
        if not np.isfinite(value):
            raise ValueError("tor_single: input is off the physical domain (I - O_S not SPD)")
        return value
    value, bound = float(res[0]), float(res[1])
    if not (np.isfinite(value) and np.isfinite(bound) and bound >= 0.0):
        raise ValueError("tor_single: off the physical domain or uncertifiable "
                         "(the bound refuses rather than overclaims)")
    return value, {"tier": tier, "abs_error_bound": bound,
                   "rel_error_bound": bound / abs(value) if value != 0 else float("inf")}


# --- Workspace (API of reference gbskernels/__init__.py:716-908) ----------------

def _square_members(matrices) -> list[np.ndarray]:
    mats = []
    for i, m in enumerate(matrices):
        a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128))
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise ValueError(f"Workspace: matrix #{i} must be square 2-D; got shape {a.shape}")
        mats.append(a)
    return mats


def _dd_takes(func: str, A: np.ndarray) -> bool:
    """Whether the double-double kernels cover this matrix (size; real O for tor)."""
    try:
        _prepare(func, A[None], "dd")
    except ValueError:
        return False
    return True


class Workspace:
    """Ragged, device-resident evaluation handle.

    A ragged call is ONE heterogeneous batch of the mixed-batch engine
    (``gbs_batch_run``): the matrices are packed into one host buffer, copied to
    the device in one transfer, and their per-size buckets run concurrently on
    a pool of HIP streams; the engine's grow-only device buffers persist across
    calls, so a sampler loop pays allocation once (:meth:`reallocs` counts the
    growths). Device-resident outputs (``*_resident``) come from a device
    session. Element ``i``'s value is identical to evaluating its matrix alone;
    ``auto`` escalates each element on its own cancellation indicator.
    ``backend="cpu"`` serves ``precision="ref"`` only."""

    def __init__(self, backend: str = "gpu") -> None:
        if backend not in ("cpu", "gpu"):
            raise ValueError(f"unknown backend {backend!r}; expected 'cpu' or 'gpu'")
        self._backend = backend
        self._engine = None    # serving.Batch, opened on first use
        self._session = None   # device session for resident outputs
        self._closed = False

    # lifecycle: handles open lazily (entering the context opens them early);
    # close() releases both and is idempotent
    def _open(self) -> None:
        if self._closed:
            raise RuntimeError("Workspace is closed")
        if self._backend == "gpu" and self._engine is None:
            from .serving import Batch
            self._engine = Batch(nstreams=4, use_graphs=False)
            self._session = _ext().Session()

    def __enter__(self) -> "Workspace":
        self._open()
        return self

    def __exit__(self, *exc) -> bool:
        self.close()
        return False

    def close(self) -> None:
        for h in (self._engine, self._session):
            if h is not None:
                h.close()
        self._engine = self._session = None
        self._closed = True

    def reallocs(self) -> int:
        """Times the engine's device buffers grew (0 before the first call)."""
        return 0 if self._engine is None else self._engine.stats()["device_reallocs"]

    def perm_batched(self, matrices, precision: str = "fp64") -> np.ndarray:
        return self._ragged("perm", matrices, precision)

    def haf_batched(self, matrices, precision: str = "fp64") -> np.ndarray:
        return self._ragged("haf", matrices, precision)

    def lhaf_batched(self, matrices, precision: str = "fp64") -> np.ndarray:
        return self._ragged("lhaf", matrices, precision)

    def tor_batched(self, matrices, precision: str = "fp64") -> np.ndarray:
        return self._ragged("tor", matrices, precision)

    def _ragged(self, func: str, matrices, precision: str) -> np.ndarray:
        if self._closed:
            raise RuntimeError("Workspace is closed")
        mats = _square_members(matrices)
        if self._backend == "cpu":
            _check_args(precision, "cpu", False, None)
            return np.array([_ref.evaluate(func, a) for a in mats], dtype=np.complex128)
        if precision not in ("fp64", "dd", "auto"):
            raise NotImplementedError(f"Workspace precision {precision!r}: use 'fp64', 'dd' or 'auto'")
        if not mats:
            return np.empty(0, dtype=np.complex128)
        run = "kappa" if precision == "auto" else precision
        # each size is checked once, against its first member (named on failure)
        first: dict[int, int] = {}
        for i, a in enumerate(mats):
            first.setdefault(a.shape[0], i)
        for d, i in first.items():
            try:
                _prepare(func, mats[i][None], run)
            except ValueError as e:
                raise ValueError(f"matrix #{i} (dim {d}): {e}") from e
        self._open()
        values, aux = self._engine.run([(func, run, a) for a in mats])
        values = values.copy()
        if precision == "auto":
            kappa = aux / np.maximum(np.abs(values), 1e-300)
            risky = np.flatnonzero(kappa >= _AUTO_KAPPA_MAX)
            dd = [i for i in risky if _dd_takes(func, mats[i])]
            if dd:
                values[dd] = self._engine.run([(func, "dd", mats[i]) for i in dd])[0]
            for i in sorted(set(risky.tolist()) - set(dd)):
                values[i] = _ref.evaluate(func, mats[i])
        return values

    def perm_resident(self, matrices, precision: str = "fp64"):
        return self._resident("perm", matrices, precision)

    def haf_resident(self, matrices, precision: str = "fp64"):
        return self._resident("haf", matrices, precision)

    def lhaf_resident(self, matrices, precision: str = "fp64"):
        return self._resident("lhaf", matrices, precision)

    def tor_resident(self, matrices, precision: str = "fp64"):
        return self._resident("tor", matrices, precision)

    def _resident(self, func: str, matrices, precision: str):
        """Uniform batch evaluated with the result left in device memory
        (a :class:`~gbskernels_amd.ext.gbskernels_ext.DeviceResult`)."""
        if self._closed:
            raise RuntimeError("Workspace is closed")
        if self._backend != "gpu":
            raise NotImplementedError("device-resident output requires backend='gpu'")
        if precision not in ("fp64", "dd"):
            raise NotImplementedError("device-resident output supports 'fp64' and 'dd'")
        try:
            stack = np.ascontiguousarray(np.asarray(matrices, dtype=np.complex128))
        except (ValueError, TypeError):
            stack = None
        if stack is None or stack.ndim != 3:
            raise ValueError("device-resident output needs a uniform (B, d, d) batch")
        stack = _prepare(func, stack, precision)
        self._open()
        return self._session._resident(func if precision == "fp64" else f"{func}_dd", stack)
