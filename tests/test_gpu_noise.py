"""Photon noise of the learned path's training pairs (fdk_poisson_noise, no
reference counterpart): counts ~ Poisson(I0 exp(-mu p)) at I0 = 1e5, then
p <- -ln(max(counts, 1) / I0) / mu.

The counts are checked against the Poisson law itself (mean, variance and a
chi-square of the histogram against scipy.stats.poisson, with the clamp of 0
to 1 folded in), against the oracle's numpy restatement of the same
counter-based sampler pixel by pixel, and for seed reproducibility."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle as O

pytestmark = pytest.mark.gpu
I0, MU = 1e5, 0.02


def noisy(torch, p_host, seed):
    from fdk_amd.unet import poisson_noise
    p = torch.as_tensor(p_host).to("cuda")
    poisson_noise(p, I0, MU, seed)
    torch.cuda.synchronize()
    return p.cpu().numpy()


def counts_of(p):
    return I0 * np.exp(-MU * p.astype(np.float64))


def clamped_pmf(lam, kmax):
    """pmf of max(K, 1), K ~ Poisson(lam), on 0..kmax (entry 0 is zero)."""
    k = np.arange(kmax + 1)
    pmf = stats.poisson.pmf(k, lam)
    pmf[1] += pmf[0]
    pmf[0] = 0.0
    return k, pmf


@pytest.mark.parametrize("lam", [5.0, 50.0, 500.0, 1e5])
def test_noise_counts_follow_the_poisson_law(torch_cuda, lam):
    n = 1 << 21
    p0 = np.full(n, -math.log(lam / I0) / MU, np.float32)
    lam = float(I0 * np.exp(-MU * np.float64(p0[0])))  # the mean the float32 input encodes
    c = counts_of(noisy(torch_cuda, p0, seed=20240))
    k = np.rint(c)
    assert np.abs(c - k).max() < 1e-3  # integer counts survive the float32 line integrals
    assert k.min() >= 1
    kk, pmf = clamped_pmf(lam, int(lam + 12 * math.sqrt(lam) + 20))
    assert pmf.sum() > 1 - 1e-12
    mean = float((kk * pmf).sum())
    var = float(((kk - mean) ** 2 * pmf).sum())
    m4 = float(((kk - mean) ** 4 * pmf).sum())
    assert abs(k.mean() - mean) < 5 * math.sqrt(var / n), (k.mean(), mean)
    assert abs(k.var() - var) < 5 * math.sqrt((m4 - var * var) / n), (k.var(), var)
    # chi-square over bins holding >= 50 expected counts each
    edges, acc = [0], 0.0
    for j in range(kk.size):
        acc += pmf[j] * n
        if acc >= 50.0:
            edges.append(j + 1)
            acc = 0.0
    edges[-1] = kk.size
    expect = np.array([pmf[a:b].sum() * n for a, b in zip(edges[:-1], edges[1:])])
    got = np.histogram(k, bins=np.array(edges, float) - 0.5)[0]
    chi2 = float(((got - expect) ** 2 / expect).sum())
    assert stats.chi2.sf(chi2, len(expect) - 1) > 1e-6, (chi2, len(expect))
    if lam <= 500:  # Poisson skew ~ 1/sqrt(lam): the normal approximation it replaces has none
        # skew of the clamped law max(K, 1) (the clamp lifts the lower tail at small lam)
        skew_law = float(((kk - mean) ** 3 * pmf).sum()) / var ** 1.5
        skew = float(stats.skew(k))
        assert abs(skew - skew_law) < 5 * math.sqrt(6.0 / n), (skew, skew_law)


def test_noise_matches_the_oracle_sampler(torch_cuda):
    """A ramp of means 1 .. 1e5 (both branches and the switch at 10) against the
    numpy restatement of the same counter-based sampler."""
    n = 1 << 18
    lam = np.geomspace(1.0, 1e5, n)
    p0 = (-np.log(lam / I0) / MU).astype(np.float32)
    got = noisy(torch_cuda, p0, seed=77)
    ref = O.poisson_noise(p0, I0, MU, 77)
    same = got == ref
    assert same.mean() > 0.9999, same.mean()
    assert np.abs(np.rint(counts_of(got)) - np.rint(counts_of(ref))).max() <= 1


def test_noise_seed_reproducible(torch_cuda):
    n = 1 << 20
    p0 = np.full(n, -math.log(500.0 / I0) / MU, np.float32)
    a = noisy(torch_cuda, p0, seed=5)
    b = noisy(torch_cuda, p0, seed=5)
    c = noisy(torch_cuda, p0, seed=6)
    assert np.array_equal(a, b)
    assert (a != c).mean() > 0.9


def test_noise_edge_cases(torch_cuda):
    import torch
    from fdk_amd.unet import poisson_noise
    empty = torch.empty(0, device="cuda")
    poisson_noise(empty, I0, MU, 1)  # no work, no error
    p = torch.tensor([float("nan"), 0.0, 1e4], device="cuda")
    poisson_noise(p, I0, MU, 1)
    torch.cuda.synchronize()
    out = p.cpu().numpy()
    assert math.isnan(out[0])  # left as it was
    assert abs(counts_of(out[1:2])[0] - I0) < 20 * math.sqrt(I0)
    assert out[2] == np.float32(-math.log(1.0 / I0) / MU)  # lambda ~ 0: zero counts clamp to 1
