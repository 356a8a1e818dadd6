"""The resident chain-rule GBS sampler: distributionally equal to the exact
sequentially truncated chain rule (computed on the host from the oracle's
quad-precision sieve), reproducible per seed, and refusing what it cannot
evaluate. The reference validates its device sampler the same way
(core/sampler_draw.cu:5-7: chi-square / TV, never bit for bit)."""
import itertools
from math import factorial

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def squeezed_interferometer_cov(M, r, seed, hbar=2.0):
    from gbskernels_amd.workloads import haar_unitary
    U = haar_unitary(M, np.random.default_rng(seed))
    O = np.block([[U.real, -U.imag], [U.imag, U.real]])
    S = O @ np.diag(np.concatenate([np.exp(-r), np.exp(r)]))
    return hbar / 2 * S @ S.T


def exact_chain_probs(oracle, cov, cutoff):
    """P(n) of the per-mode truncated chain rule for every n in [0..cutoff]^M."""
    from gbskernels_amd.sampling import reduced_A
    M = cov.shape[0] // 2
    C = cutoff
    pats = list(itertools.product(range(C + 1), repeat=M))
    cond = {}
    for k in range(1, M + 1):
        A = reduced_A(cov, k)
        prefixes = sorted({p[:k - 1] for p in pats})
        reps = []
        for pre in prefixes:
            for j in range(C + 1):
                reps.append(list(pre) + [j] + list(pre) + [j])
        h = oracle.lhaf_repeated(A, None, np.array(reps, np.int32), quad=True).real
        w = np.clip(h, 0, None).reshape(len(prefixes), C + 1) / np.array([factorial(j) for j in range(C + 1)])
        for i, pre in enumerate(prefixes):
            cond[(k, pre)] = w[i] / w[i].sum()
    prob = {}
    for p in pats:
        v = 1.0
        for k in range(1, M + 1):
            v *= cond[(k, p[:k - 1])][p[k - 1]]
        prob[p] = v
    return prob


def test_distribution_matches_exact_chain(oracle):
    from scipy.stats import chi2
    from gbskernels_amd.sampling import sample_resident
    M, C, D = 3, 4, 200000
    cov = squeezed_interferometer_cov(M, np.array([0.7, 0.5, 0.3]), 11)
    prob = exact_chain_probs(oracle, cov, C)
    assert abs(sum(prob.values()) - 1.0) < 1e-12
    s = sample_resident(cov, D, cutoff=C, seed=7)
    assert s.shape == (D, M) and s.min() >= 0 and s.max() <= C
    keys, counts = np.unique(s, axis=0, return_counts=True)
    emp = {tuple(int(x) for x in k): c for k, c in zip(keys, counts)}
    exp = np.array([prob[p] * D for p in prob])
    obs = np.array([emp.get(p, 0) for p in prob])
    big = exp >= 5
    stat = float(np.sum((obs[big] - exp[big]) ** 2 / exp[big]))
    rest_o, rest_e = obs[~big].sum(), exp[~big].sum()
    dof = int(big.sum())
    if rest_e > 0:
        stat += (rest_o - rest_e) ** 2 / rest_e
    else:
        dof -= 1
    pval = float(chi2.sf(stat, dof))
    tv = 0.5 * float(np.abs(obs / D - exp / D).sum())
    print(f"chi2 {stat:.1f} dof {dof} p {pval:.3g} tv {tv:.4f}", flush=True)
    assert pval > 1e-6 and tv < 0.01


def test_reproducible_and_edges():
    from gbskernels_amd.ext import gbskernels_ext as ext
    from gbskernels_amd.sampling import reduced_A, sample_resident
    cov = squeezed_interferometer_cov(4, np.array([0.6, 0.4, 0.5, 0.3]), 3)
    a = sample_resident(cov, 5000, cutoff=5, seed=123)
    b = sample_resident(cov, 5000, cutoff=5, seed=123)
    c = sample_resident(cov, 5000, cutoff=5, seed=124)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.all(sample_resident(cov, 100, cutoff=0, seed=1) == 0)
    assert sample_resident(cov, 0, cutoff=3).shape == (0, 4)
    # the first mode's marginal is the same whatever the seed's other draws
    assert np.array_equal(sample_resident(cov, 64, cutoff=5, seed=9)[:, 0],
                          sample_resident(cov, 128, cutoff=5, seed=9)[:64, 0])
    with pytest.raises(ValueError):   # 17 modes: A_17 is 34 x 34, past the sieve's side
        ext.sample_resident(np.zeros(10), np.zeros(17, np.int32), 17, 4, 2, 0, np.ones(3), 0)
    # a conditional past the sieve's domain (N = 2 j > 64 at j = 33) fails the call
    cov1 = squeezed_interferometer_cov(1, np.array([0.8]), 5)
    with pytest.raises(ValueError):
        sample_resident(cov1, 16, cutoff=33, seed=3)
    assert sample_resident(cov1, 16, cutoff=32, seed=3).shape == (16, 1)
