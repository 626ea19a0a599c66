"""The oracle's photon-noise sampler (a numpy restatement of the library's
counter-based Poisson sampler) pinned on the CPU: its uniforms against a
pure-Python integer restatement of the hash, and its counts against the Poisson
law (scipy.stats.poisson) in both branches (Knuth below 10 expected counts,
Hoermann's PTRS from 10 up)."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle as O

M32 = 0xFFFFFFFF


def hash32_py(x):
    x &= M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def unif_py(seed, i, k):
    hi = ((i * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)) >> 32
    a = hash32_py(seed ^ hi)
    c = hash32_py((i & M32) + ((0x632BE5AB * k) & M32))
    return (hash32_py(a ^ c) + 0.5) / 4294967296.0


def test_uniforms_match_integer_restatement():
    idx = np.array([0, 1, 2, 12345, 2 ** 31 + 7, 2 ** 40 + 3], np.uint64)
    for seed in (0, 77, 0xDEADBEEF):
        for k in (0, 1, 5, 255):
            got = O._noise_unif(seed, idx, k)
            want = [unif_py(seed, int(i), k) for i in idx]
            assert np.array_equal(got, np.array(want))


@pytest.mark.parametrize("lam", [3.0, 9.5, 50.0, 1e5])
def test_counts_follow_the_poisson_law(lam):
    n = 200_000
    k = O.poisson_counts(np.full(n, lam), seed=11)
    assert np.all(k == np.rint(k)) and k.min() >= 0
    sd = math.sqrt(lam)
    assert abs(k.mean() - lam) < 5 * sd / math.sqrt(n)
    assert abs(k.var() - lam) < 5 * math.sqrt((lam * (1 + 2 * lam)) / n)
    lo, hi = int(max(0, lam - 6 * sd)), int(lam + 6 * sd) + 1
    kk = np.arange(lo, hi + 1)
    pmf = stats.poisson.pmf(kk, lam)
    edges, acc = [lo], 0.0
    for j, pj in enumerate(pmf):
        acc += pj * n
        if acc >= 50.0:
            edges.append(kk[j] + 1)
            acc = 0.0
    edges[-1] = hi + 1
    expect = np.array([stats.poisson.cdf(b - 1, lam) - stats.poisson.cdf(a - 1, lam)
                       for a, b in zip(edges[:-1], edges[1:])]) * n
    got = np.histogram(k, bins=np.array(edges, float) - 0.5)[0]
    chi2 = float(((got - expect) ** 2 / expect).sum())
    assert stats.chi2.sf(chi2, len(expect) - 1) > 1e-6, chi2


def test_counts_reproducible_and_seeded():
    lam = np.geomspace(1.0, 1e5, 4096)
    a = O.poisson_counts(lam, seed=5)
    b = O.poisson_counts(lam, seed=5)
    c = O.poisson_counts(lam, seed=6)
    assert np.array_equal(a, b)
    assert (a != c).mean() > 0.5
