"""GBS probability layer (reference sampling/gbs.py) on the GPU kernels: photon
probabilities against brute-force matchings of the expanded pattern matrices,
normalisation, the displaced (loop-hafnian) path, and threshold probabilities
two ways (marginalised photon numbers, torontonians) that must agree."""
from math import cosh, factorial, prod

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def pure_state(m, seed, r_lo=0.1, r_hi=0.45):
    """(B, r, cov) of one squeezed-then-interfered pure state (xxpp, hbar = 2)."""
    g = np.random.default_rng(seed)
    r = g.uniform(r_lo, r_hi, m)
    z = (g.standard_normal((m, m)) + 1j * g.standard_normal((m, m))) / np.sqrt(2.0)
    q, rr = np.linalg.qr(z)
    ph = np.diagonal(rr).copy()
    U = q * (ph / np.abs(ph))
    B = U @ np.diag(np.tanh(r)) @ U.T
    S = np.block([[U.real, -U.imag], [U.imag, U.real]]) @ np.diag(np.concatenate([np.exp(-r), np.exp(r)]))
    return B, r, S @ S.T


def test_probabilities_match_brute_force(oracle):
    from gbskernels_amd import gbs
    B, r, _ = pure_state(3, 4)
    pats, probs = gbs.probabilities(B, r, 4)
    norm = prod(cosh(x) for x in r)
    for n, p in zip(pats, probs):
        E = oracle.expand_repeated(B, None, n)
        h = oracle.matchings(E)[0] if E.shape[0] else 1.0
        want = abs(h) ** 2 / (prod(factorial(x) for x in n) * norm)
        assert abs(p - want) <= 1e-13 + 1e-11 * want, n
    assert 0.99 < gbs.total_probability(B, r, 10) <= 1.0 + 1e-12


def test_displaced_path(oracle):
    from gbskernels_amd import gbs
    B, r, _ = pure_state(3, 8)
    _, p0 = gbs.probabilities(B, r, 4)
    _, pd0 = gbs.displaced_probabilities(B, r, np.zeros(3), 4)
    assert np.allclose(pd0, p0, rtol=1e-12, atol=1e-15)
    alpha = np.array([0.2 + 0.1j, -0.15j, 0.1])
    pats, pd = gbs.displaced_probabilities(B, r, alpha, 4)
    A = -B
    gamma = alpha - A @ np.conj(alpha)
    pref = np.exp(-0.5 * (np.vdot(alpha, alpha) - alpha.conj() @ A @ alpha.conj()))
    norm = prod(cosh(x) for x in r)
    for n, p in list(zip(pats, pd))[::7]:
        E = oracle.expand_repeated(A, gamma, n)
        lh = oracle.matchings(E, loop=True)[0] if E.shape[0] else 1.0
        want = abs(pref * lh) ** 2 / (prod(factorial(x) for x in n) * norm)
        assert abs(p - want) <= 1e-13 + 1e-11 * want, n
    assert 0.99 < gbs.displaced_total(B, r, alpha, 9) <= 1.0 + 1e-12


def test_threshold_two_ways():
    import gbskernels_amd as gk
    from gbskernels_amd import gbs
    B, r, cov = pure_state(3, 5, 0.2, 0.4)
    direct = gbs.torontonian_threshold_probabilities(cov)
    assert abs(sum(direct.values()) - 1.0) < 1e-12            # exact distribution, no cutoff
    marg = gbs.threshold_probabilities(B, r, 12)
    for c, p in direct.items():
        assert abs(p - marg[c]) < 1e-6, c                      # the marginal's cutoff tail
    # the complex-basis O = I - Q^-1 (the reference's route, Hermitian: the LU
    # fallback) gives the same probabilities
    from gbskernels_amd.ext import gbskernels_ext as ext
    from gbskernels_amd.sampling import qmat
    Q = qmat(cov)
    O = np.eye(6) - np.linalg.inv(Q)
    full = ext.tor(O[None])[0] / np.sqrt(np.linalg.det(Q)).real
    assert abs(full - direct[(1, 1, 1)]) < 1e-12
    # the real xxpp matrix gives the same click probabilities through tor_single
    Ox, lsd = gbs.threshold_O_xxpp(cov)
    m = 3
    for c in [(1, 0, 0), (1, 1, 0), (1, 1, 1)]:
        S = [i for i in range(m) if c[i]]
        idx = S + [i + m for i in S]
        p = gk.tor_single(np.ascontiguousarray(Ox[np.ix_(idx, idx)])) * np.exp(-lsd)
        assert abs(p - direct[c]) < 1e-12, c


def test_contract():
    from gbskernels_amd import gbs
    B, r, _ = pure_state(2, 1)
    with pytest.raises(NotImplementedError):
        gbs.probabilities(B, r, 3, backend="cpu")
    with pytest.raises(NotImplementedError):
        gbs.probabilities(B, r, 3, precision="dd")
    assert len(gbs.fock_patterns(3, 2)) == 8
    Bk, rk = gbs.random_gbs_kernel(4, 0)
    assert np.allclose(Bk, Bk.T) and np.all((rk >= 0.1) & (rk <= 0.45))
