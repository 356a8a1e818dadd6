"""The ``"ref"`` precision tier: arbitrary-precision (mpmath) evaluation.

The reference exposes this tier (reference ``gbskernels/__init__.py:54-63``,
``highprec_ref/``) as the CPU ground truth, and its GPU ``auto`` and
``certified`` ladders escalate to it explicitly (reference
``gbskernels/__init__.py:248-269,421-429``); results it produces are labelled
``tier="ref"`` in diagnostics. It is exponential-time by definition and meant
for small sizes only. Algorithms (this module's own code):

* permanent: Ryser inclusion-exclusion over column subsets;
* hafnian / loop hafnian: recursion over (loop) perfect matchings;
* torontonian: subset sum of ``1/sqrt(det(I - O_S))`` with mpmath determinants.

Input entries are converted exactly (binary64 -> mpf is exact).
"""

from __future__ import annotations

import itertools

import numpy as np


def _mp():
    import mpmath
    return mpmath


def _mat(A):
    M = np.asarray(A, dtype=np.complex128)
    if M.ndim != 2 or M.shape[0] != M.shape[1]:
        raise ValueError(f"expected a square matrix, got {M.shape}")
    return M


def perm(A, dps: int = 60):
    mp = _mp()
    M = _mat(A)
    n = M.shape[0]
    with mp.workdps(dps):
        if n == 0:
            return mp.mpc(1)
        E = [[mp.mpc(complex(M[i, j])) for j in range(n)] for i in range(n)]
        total = mp.mpc(0)
        for k in range(1, n + 1):
            for S in itertools.combinations(range(n), k):
                p = mp.mpc(1)
                for i in range(n):
                    p *= mp.fsum(E[i][j] for j in S)
                total += (-1) ** k * p
        return (-1) ** n * total


def _matchings(E, idx, loop):
    mp = _mp()
    if not idx:
        return mp.mpc(1)
    i, rest = idx[0], idx[1:]
    total = E[i][i] * _matchings(E, rest, loop) if loop else mp.mpc(0)
    for k, j in enumerate(rest):
        total += E[i][j] * _matchings(E, rest[:k] + rest[k + 1:], loop)
    return total


def haf(A, dps: int = 60):
    mp = _mp()
    M = _mat(A)
    N = M.shape[0]
    with mp.workdps(dps):
        if N % 2:
            return mp.mpc(0)
        E = [[mp.mpc(complex(M[i, j])) for j in range(N)] for i in range(N)]
        return _matchings(E, tuple(range(N)), False)


def lhaf(A, dps: int = 60):
    mp = _mp()
    M = _mat(A)
    N = M.shape[0]
    with mp.workdps(dps):
        E = [[mp.mpc(complex(M[i, j])) for j in range(N)] for i in range(N)]
        return _matchings(E, tuple(range(N)), True)


def tor(O, dps: int = 60):
    mp = _mp()
    M = _mat(O)
    if M.shape[0] % 2:
        raise ValueError("torontonian needs an even dimension")
    n = M.shape[0] // 2
    with mp.workdps(dps):
        total = mp.mpc(0)
        for m in range(n + 1):
            for S in itertools.combinations(range(n), m):
                idx = list(S) + [i + n for i in S]
                if m == 0:
                    det = mp.mpc(1)
                else:
                    sub = mp.matrix(2 * m, 2 * m)
                    for r, a in enumerate(idx):
                        for c, b in enumerate(idx):
                            sub[r, c] = (1 if r == c else 0) - mp.mpc(complex(M[a, b]))
                    det = mp.det(sub)
                total += (-1) ** (n - m) / mp.sqrt(det)
        return total


REF = {"perm": perm, "haf": haf, "lhaf": lhaf, "tor": tor}


def evaluate(func: str, A) -> complex:
    return complex(REF[func](A))
