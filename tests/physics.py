"""Synthetic physical inputs (Gaussian-state constructions written for these
tests; numpy only). Threshold-detector O matrices follow the definition the
reference uses (sampling/gbs.py:241-275: O = I - Sigma_x^-1 with
Sigma_x = (cov + I) / 2, hbar = 2), built here from our own state model."""
import numpy as np


def haar_unitary(m, rng):
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    return q * (np.diagonal(r) / np.abs(np.diagonal(r)))


def threshold_O(nm, rng, eta=0.7, rmin=0.3, rmax=1.0):
    """Real 2nm x 2nm O (xxpp) of two-mode squeezed pairs (modes 2i, 2i+1,
    r ~ U[rmin, rmax]) through a Haar interferometer with loss eta."""
    S = np.eye(2 * nm)
    for i in range(0, nm - 1, 2):
        r = rng.uniform(rmin, rmax)
        ch, sh = np.cosh(r), np.sinh(r)
        a, b = i, i + 1
        T = np.eye(2 * nm)
        T[a, a] = T[b, b] = ch
        T[a, b] = T[b, a] = sh                      # x quadratures
        T[nm + a, nm + a] = T[nm + b, nm + b] = ch
        T[nm + a, nm + b] = T[nm + b, nm + a] = -sh  # p quadratures
        S = T @ S
    U = haar_unitary(nm, rng)
    SU = np.block([[U.real, -U.imag], [U.imag, U.real]])
    S = SU @ S
    cov = eta * (S @ S.T) + (1.0 - eta) * np.eye(2 * nm)
    sig = (cov + np.eye(2 * nm)) / 2.0
    sig = (sig + sig.T) / 2.0
    O = np.eye(2 * nm) - np.linalg.inv(sig)
    return np.ascontiguousarray((O + O.T) / 2.0)
