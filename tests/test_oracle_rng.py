"""Oracle RNG pins: Philox4x32-10 KATs, seed-mix KATs, Ziggurat statistics (CPU)."""

import numpy as np
import pytest
from scipy import stats

from oracle import cpu_baseline, philox

# Random123 known-answer vectors for philox4x32_10 (counter, key, result).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    (
        (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
        (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1),
    ),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(ctr, key, expect):
    out = philox.philox4x32_10(np.array(ctr, dtype=np.uint32), key)
    assert tuple(int(x) for x in out) == expect


def test_philox_vectorised_matches_scalar():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, size=(64, 4), dtype=np.uint64).astype(np.uint32)
    batch = philox.philox4x32_10(ctr, (123, 456))
    for i in range(0, 64, 9):
        assert np.array_equal(batch[i], philox.philox4x32_10(ctr[i], (123, 456)))


def test_box_muller_is_standard_normal():
    ctr = np.zeros((50000, 4), np.uint32)
    ctr[:, 0] = np.arange(50000)
    z = philox.normals4(ctr, seed=7).reshape(-1)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3


def test_seed_mix_matches_splitmix64_outputs():
    # ssms_rng_mix_seed(base, k, 0) with base=0 is SplitMix64's k-th output for seed 0
    # (gsl_rng.h:97-102; Steele, Lea & Flood 2014): 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, ...
    assert cpu_baseline.mix_seed(0, 1, 0) == 0xE220A8397B1DCDAF
    assert cpu_baseline.mix_seed(0, 2, 0) == 0x6E789E6AA1B965F4
    assert cpu_baseline.mix_seed(0, 3, 0) == 0x06C45D188009454F


def test_seed_mix_unique_over_grid():
    seeds = {cpu_baseline.mix_seed(42, k, n) for k in range(50) for n in range(200)}
    assert len(seeds) == 50 * 200


def test_ziggurat_normals_statistics():
    # moments / KS / tail mass, as tests/test_ziggurat_validation.py checks the GSL Ziggurat
    z = np.concatenate([cpu_baseline.ziggurat_normals(3, k, 0, 20000) for k in range(10)]).astype(np.float64)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.015
    assert abs(stats.skew(z)) < 0.03
    assert abs(stats.kurtosis(z)) < 0.06
    assert stats.kstest(z, "norm").pvalue > 1e-3
    for k in (3.0, 3.5):
        expect = 2 * stats.norm.sf(k)
        got = np.mean(np.abs(z) > k)
        assert abs(got - expect) < 5 * np.sqrt(expect / z.size) + 1e-5


def test_box_muller_angle_bits():
    # the angle is the top 23 bits in revolutions: 0 for u < 2^9, one step 2^-23 at u = 2^9,
    # 1 - 2^-23 for the largest u, and 1/2 at the top bit alone
    u = np.array([0, 511, 512, 0xFFFFFFFF, 0x80000000], np.uint32)
    assert philox.rev(u).tolist() == [0.0, 0.0, 2.0**-23, 1.0 - 2.0**-23, 0.5]
    # s scales both normals of a pair; negative s gives the same law (|s|)
    a = np.array([123456789, 987654321], np.uint32)
    b = np.array([555555555, 3000000000], np.uint32)
    z0, z1 = philox.box_muller(a, b)
    w0, w1 = philox.box_muller(a, b, s=-2.5)
    assert np.allclose(w0, 2.5 * z0, rtol=1e-15) and np.allclose(w1, 2.5 * z1, rtol=1e-15)
