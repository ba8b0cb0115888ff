"""Pins of the KDE oracle (CPU): scikit-learn, the reference's bandwidth value, closed forms."""

import math

import numpy as np
import pytest

from oracle import cpu_baseline, kde_ref

sklearn = pytest.importorskip("sklearn.neighbors")


def test_bandwidth_matches_reference_test_value():
    # tests/test_kde_class.py:31-36: constant sample -> ~0.0008 with std_cutoff 1e-3
    bw = kde_ref.bandwidth_silverman(np.array([1.0, 1.0, 1.0]))
    assert abs(bw - 0.0008) < 0.001 and bw > 0
    assert abs(bw - (4 / 9) ** 0.2 * 1e-3) < 1e-15
    # n == 1 -> std = 10
    assert abs(kde_ref.bandwidth_silverman([2.0]) - (4 / 3) ** 0.2 * 10) < 1e-12


def test_logsum_matches_sklearn_score_samples():
    rng = np.random.default_rng(0)
    data = rng.normal(0.3, 0.5, 700)
    x = rng.uniform(-2, 3, 50)
    h = kde_ref.bandwidth_silverman(data)
    kde = sklearn.KernelDensity(kernel="gaussian", bandwidth=h).fit(data[:, None])
    ref = kde.score_samples(x[:, None])
    ours = kde_ref.logsumexp_gauss(x, data, h) - math.log(data.size) - math.log(h) - 0.5 * math.log(2 * math.pi)
    assert np.max(np.abs(ours - ref)) < 1e-9


def test_cpu_logsum_matches_numpy():
    rng = np.random.default_rng(1)
    s = rng.normal(size=3000)
    x = rng.normal(size=200) * 3
    a = cpu_baseline.kde_logsum(x, s, 0.1)
    e = -((x[:, None] - s[None, :]) ** 2) / (2 * 0.01)
    m = e.max(1)
    b = m + np.log(np.exp(e - m[:, None]).sum(1))
    assert np.max(np.abs(a - b)) < 1e-9


def test_single_sample_kde_is_a_lognormal_density():
    # one sample per choice: density of log rt is N(log r0, h) with h = (4/3)^(1/5)*10
    fit = kde_ref.KDEFit(np.array([1.5, 0.7], np.float32), np.array([1, -1]), [-1, 1])
    rt = np.array([1.2])
    h = (4 / 3) ** 0.2 * 10
    x = np.log(1.2)
    expect = math.log(0.5) - 0.5 * ((x - math.log(np.float32(1.5))) / h) ** 2 - math.log(h) - 0.5 * math.log(2 * math.pi) - x
    assert abs(fit.eval(rt, np.array([1]))[0] - expect) < 1e-9


def test_floor_and_no_base_data():
    rts = np.array([0.5, 0.6, 0.7, -999.0], np.float32)
    ch = np.array([1, 1, 1, 1])
    fit = kde_ref.KDEFit(rts, ch, [-1, 1])
    out = fit.eval(np.array([0.6, -0.1, 15.0, 0.6]), np.array([1, 1, 1, -1]), max_t=20.0)
    assert out[1] == -66.774  # rt <= 0
    assert out[2] == -66.774  # far tail below the floor
    assert abs(out[3] - (-math.log(4 * 20.0))) < 1e-12  # choice without base data
    assert fit.p[1] == 1.0 and fit.logs[1].size == 3  # omission counted in p, dropped from the data


def test_sample_counts_round_the_float64_product():
    # kde_class.py:387-399: p_c = count / N as a Python float, round(n p_c) on the float64 product.
    # Where n count / N is exactly k + 1/2 in rationals, the float64 product can land just below
    # or above it; then the count follows the float, not the rational. Checked against the
    # float64 product formed from exact rationals (Fraction) and rounded by hand.
    from fractions import Fraction as Q

    def by_hand(n, c, N):
        p = Q(c / N)  # the float64 p_c, exactly
        prod = Q(float(n * p))  # n p_c rounded once to float64 (n * p exact, then float())
        f = prod.numerator // prod.denominator
        rem = prod - f
        return f + (1 if rem > Q(1, 2) or (rem == Q(1, 2) and f % 2 == 1) else 0)

    cases = [(320, 23), (320, 49), (320, 51), (1600, 1), (7, 3), (4000, 2999)]
    for N, c in cases:
        hand = [by_hand(800, c, N), by_hand(800, N - c, N)]
        first_max = int(np.argmax(hand))
        hand[first_max] += 800 - sum(hand)  # the +-1 fix-up (the sum is off by at most one)
        assert kde_ref.n_by_choice(800, [c / N, (N - c) / N]) == hand, (N, c)
    # the rational half-to-even rule differs here: it would give 58 + 742 and 122 + 678
    assert kde_ref.n_by_choice(800, [23 / 320, 297 / 320]) == [57, 743]
    assert kde_ref.n_by_choice(800, [49 / 320, 271 / 320]) == [123, 677]
    # counts that add up to more than n: the first largest loses one (np.argmax)
    assert kde_ref.n_by_choice(3, [0.5, 0.5]) == [1, 2]
