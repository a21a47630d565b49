"""Statistical pin of the per-row connection generator.

The reference draws every connection from one sequential stream
(network.cpp:154-204): each ordered (pre, post) pair, autapses excluded, is
an independent Bernoulli(p_primary) trial; each primary excitatory synapse
gets a tripartite astrocyte with probability p_third_if_primary, drawn from
the postsynaptic neuron's pool. The per-row generator (builder.cpp,
"connectivity.generator": "per_row") draws the same distribution from one
counter stream per presynaptic row, by geometric gaps. Its bits differ from
the reference's, so it is pinned here by that distribution: counts within
binomial bounds, out- and in-degree histograms against the binomial law,
targets uniform over the columns, tripartite counts and pools, and
agreement with the sequential (reference-exact) generator's statistics.
"""

import math

import numpy as np
import pytest
from scipy import stats

from conftest import config_path

import astrosimgpu_amd as A

N_E, N_I, N_A = 1600, 400, 400
P, P3 = 0.05, 0.2


def _model(generator, seed):
    return A.Model(config_path("quick"), seed=seed, assignments={
        "N.N_E": N_E, "N.N_I": N_I, "N.N_A": N_A, "connectivity.p_primary": P,
        "connectivity.p_third_if_primary": P3, "connectivity.generator": generator})


def _rows(m, name):
    row, tgt, _, _ = m.csr(name)
    return np.asarray(row, np.int64), np.asarray(tgt, np.int64)


def _within(count, trials, p, sigmas=5.0):
    mean = trials * p
    sd = math.sqrt(trials * p * (1.0 - p))
    return abs(count - mean) < sigmas * sd, (count, mean, sd)


def _binomial_gof(degrees, n, p):
    """Chi-square of a degree histogram against Binomial(n, p), bins with
    expected count >= 5 (the tails pooled)."""
    k = np.arange(n + 1)
    pmf = stats.binom.pmf(k, n, p)
    obs = np.bincount(degrees, minlength=n + 1)[: n + 1]
    exp = pmf * len(degrees)
    keep = exp >= 5
    lo, hi = np.argmax(keep), len(keep) - np.argmax(keep[::-1])
    o = np.concatenate([[obs[:lo].sum()], obs[lo:hi], [obs[hi:].sum()]])
    e = np.concatenate([[exp[:lo].sum()], exp[lo:hi], [exp[hi:].sum()]])
    mask = e > 0
    chi2 = float((((o - e) ** 2)[mask] / e[mask]).sum())
    return stats.chi2.sf(chi2, mask.sum() - 1)


@pytest.mark.parametrize("seed", [1, 2])
def test_primary_counts_and_degrees_follow_the_binomial_law(seed):
    m = _model("per_row", seed)
    n = N_E + N_I
    for name, pre in (("exc", N_E), ("inh", N_I)):
        row, tgt = _rows(m, name)
        ok, info = _within(len(tgt), pre * (n - 1), P)
        assert ok, (name, info)
        out_deg = np.diff(row)
        assert _binomial_gof(out_deg, n - 1, P) > 1e-4, name
        # each row: targets strictly ascending, no autapse
        for r in range(pre):
            t = tgt[row[r]:row[r + 1]]
            assert np.all(np.diff(t) > 0)
            src = r if name == "exc" else N_E + r
            assert not np.any(t == src)
    # in-degree of every neuron from the excitatory population
    row, tgt = _rows(m, "exc")
    in_deg = np.bincount(tgt, minlength=n)
    assert _binomial_gof(in_deg[N_E:], N_E, P) > 1e-4          # inhibitory targets: N_E sources
    assert _binomial_gof(in_deg[:N_E], N_E - 1, P) > 1e-4      # excitatory: no autapse


@pytest.mark.parametrize("seed", [3])
def test_targets_are_uniform_over_the_columns(seed):
    """Geometric gaps that drifted or restarted badly would bend this."""
    m = _model("per_row", seed)
    _, tgt = _rows(m, "exc")
    hist = np.bincount(tgt * 20 // (N_E + N_I), minlength=20)
    assert stats.chisquare(hist).pvalue > 1e-4
    # position within the row is not biased either: the first target of a
    # row follows the geometric law of the first success
    row, _ = _rows(m, "exc")
    first = np.array([tgt[row[r]] for r in range(N_E) if row[r + 1] > row[r]])
    assert abs(first.mean() - (1 - P) / P) < 6 * math.sqrt((1 - P) / P ** 2 / len(first)) + 1.0


def test_tripartite_counts_and_pools():
    m = _model("per_row", 4)
    _, exc = _rows(m, "exc")
    n2a_row, n2a = _rows(m, "n2a")
    a2n_row, a2n = _rows(m, "a2n")
    ok, info = _within(len(n2a), len(exc), P3)
    assert ok, info
    assert len(n2a) == len(a2n)
    # a2n: astrocyte -> neuron, the astrocyte from the neuron's block pool
    # (pool size 1 in quick.json: astrocyte floor(local * N_A / count))
    n = N_E + N_I
    for a in range(N_A):
        posts = a2n[a2n_row[a]:a2n_row[a + 1]]
        local = np.where(posts < N_E, posts, posts - N_E)
        count = np.where(posts < N_E, N_E, N_I)
        assert np.all(local * N_A // count == a)


def test_per_row_statistics_match_the_sequential_generator():
    """Same distribution as the reference's sequential stream: mean
    out-degree, in-degree spread and tripartite fraction agree within their
    sampling error."""
    seq = _model("sequential", 5)
    per = _model("per_row", 5)
    for name in ("exc", "inh", "n2a"):
        a = len(_rows(seq, name)[1])
        b = len(_rows(per, name)[1])
        assert abs(a - b) < 5 * math.sqrt(a + b), (name, a, b)
    n = N_E + N_I
    in_a = np.bincount(_rows(seq, "exc")[1], minlength=n)
    in_b = np.bincount(_rows(per, "exc")[1], minlength=n)
    assert stats.ks_2samp(in_a, in_b).pvalue > 1e-4
