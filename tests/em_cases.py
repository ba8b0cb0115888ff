"""Shared inputs for the EM tests: exact float32 rounding and start-point-sensitive cases."""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
from scipy import stats

from oracle import em_numpy as E

F32 = np.float32


def round32(q: Fraction) -> np.float32:
    """Round an exact rational to the nearest float32, ties to even (IEEE 754 binary32,
    normal and subnormal range), by integer arithmetic only."""
    if q == 0:
        return F32(0.0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    e = max(e, -126)  # subnormals share the smallest normal exponent's quantum
    scaled = q * Fraction(2) ** (23 - e)
    m, rem = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * rem
    if twice > scaled.denominator or (twice == scaled.denominator and m % 2 == 1):
        m += 1
    return F32(sign * float(Fraction(m) * Fraction(2) ** (e - 23)))


def start_sensitive_ddm_cases(n_search: int = 4000, seed: int = 5):
    """Angle-bound DDM parameter sets whose walk length depends on how the start point
    -b0 + 2 z b0 is rounded: the float form of ddm_models.pyx:709 or the double-literal form
    of :2305 (full_ddm with zero variability).

    With v = 0 and s = 0 every increment is exactly 0, so y stays at the start y0 and the walk
    ends at the first boundary index with b[ix] < |y0|. z lies within 6e-4 of 0 or 1, so
    |y0| is close to b0, and theta is small enough that b(t) = a - t tan(theta) falls by about
    half a float32 ulp per step: one ulp of y0 moves the crossing by about two steps.

    Returns (v, a, z, theta, y0) float32 arrays of the sets where the two oracle walks end on
    different steps, y0 being the float start."""
    rng = np.random.default_rng(seed)
    n = n_search
    a = rng.uniform(1.0, 1.99, n).astype(F32)
    near = rng.uniform(1e-5, 5e-4, n) / a
    z = np.where(rng.random(n) < 0.5, 1.0 - near, near).astype(F32)
    theta = rng.uniform(3e-5, 6e-5, n).astype(F32)
    v = np.zeros(n, F32)
    zeros = np.zeros(n, F32)
    dl = np.full(n, 999, F32)
    common = dict(n_samples=1, delta_t=0.001, boundary="angle", bparams={"a": a, "theta": theta},
                  normals=np.zeros((n, 1), F32))
    _, _, s0 = E.ddm_flexbound(v, a, z, zeros, dl, zeros, **common)
    _, _, s1 = E.ddm_flexbound(v, a, z, zeros, dl, zeros, sz=zeros, sv=zeros, st=zeros,
                               uniforms=np.full((n, 3), 0.5, F32), **common)
    d = s0 != s1
    return v[d], a[d], z[d], theta[d], E.ddm_start32(z[d], a[d])


def quantile_check(x, y, qs=(0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9)):
    """The q-quantiles of two samples agree within 3 Monte-Carlo standard errors (order-statistic
    intervals of +-sqrt(q(1-q)/n) in probability)."""
    for q in qs:
        def se(s):
            d = math.sqrt(q * (1 - q) / s.size)
            lo, hi = np.quantile(s, [max(q - d, 0), min(q + d, 1)])
            return (hi - lo) / 2

        tol = 3 * math.hypot(se(x), se(y))
        assert abs(np.quantile(x, q) - np.quantile(y, q)) <= tol + 1e-6, (q, np.quantile(x, q), np.quantile(y, q), tol)


def compare_samples(r_gpu, c_gpu, r_cpu, c_cpu, choices):
    """Per choice: probabilities within 3 MC SE; with >= 200 draws each, two-sample KS p > 0.01 and
    the 0.1..0.9 rt quantiles within 3 MC SE (tests/test_parallel_equivalence.py:39-108 of the
    reference uses the same criteria)."""
    n1, n2 = r_gpu.size, r_cpu.size
    for c in choices:
        p1, p2 = np.mean(c_gpu == c), np.mean(c_cpu == c)
        se = math.sqrt(p1 * (1 - p1) / n1 + p2 * (1 - p2) / n2)
        assert abs(p1 - p2) <= 3 * se + 1e-9, (c, p1, p2, se)
        x, y = r_gpu[c_gpu == c], r_cpu[c_cpu == c]
        if min(x.size, y.size) >= 200:
            assert stats.ks_2samp(x, y).pvalue > 0.01, c
            quantile_check(x, y)


def lba_float32_deviation_ok(r32, c32, r64, c64, v, sd, t, uniforms, normals, n_samples, tan32=None, tan64=None,
                             a=None, z=None):
    """Whether float32 LBA results (the kernel, em_numpy.lba) stay within the rounding error of a
    float32 evaluation of the reference's float64 form (em_numpy.lba64) on the same draws.

    x = (a - u z) / (|v + sd g| + tan theta), each float32 operation rounding with unit roundoff
    e = 2^-24: the numerator is off by at most e (|u z| + |a - u z|), the denominator by
    e (|sd g| + |v + sd g| + |den|) + |tan32 - tan64|, the quotient by e more, so x carries the
    relative error num_err / |num| + den_err / |den| + e, which grows where the numerator or the
    denominator cancels; rt = x + t adds e |rt|. Checked with a factor 2 of slack. Returns the
    share of differing choices and whether every rt of a matching choice lies within its bound."""
    K, P = np.asarray(v).shape
    ls = np.repeat(np.arange(K), int(n_samples))
    f8 = np.float64
    e = 2.0 ** -24
    v64 = np.asarray(v, f8)[ls]
    sdg = np.broadcast_to(np.asarray(sd, f8).reshape(K, -1), (K, P))[ls] * np.asarray(normals, f8)
    vs = np.abs(v64 + sdg)
    den = vs.copy()
    dtan = np.zeros_like(den)
    if tan64 is not None:
        t64 = np.asarray(tan64, f8).reshape(-1)[ls, None]
        den = den + t64
        dtan = np.abs(np.asarray(tan32, np.float32).reshape(-1).astype(f8)[ls, None] - t64)
    a64 = np.asarray(a, f8).reshape(K, -1)[:, 0][ls, None]
    zs = np.asarray(uniforms, f8) * np.asarray(z, f8).reshape(K, -1)[:, 0][ls, None]
    num = a64 - zs
    rel = (e * (np.abs(zs) + np.abs(num)) / np.abs(num) + (e * (np.abs(sdg) + vs + np.abs(den)) + dtan) / np.abs(den)
           + e)
    j = np.asarray(c64, np.int64)
    rows = np.arange(j.size)
    x64 = num[rows, j] / den[rows, j]
    bound = 2 * (np.abs(x64) * rel[rows, j] + e * np.abs(np.asarray(r64, f8)))
    ok = (np.asarray(c32) == j) & (np.asarray(r32) != -999.0) & (np.asarray(r64) != -999.0)
    err = np.abs(np.asarray(r32, f8) - np.asarray(r64, f8))
    return float(np.mean(np.asarray(c32) != j)), bool(np.all(err[ok] <= bound[ok]))
