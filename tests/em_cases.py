"""Shared inputs for the EM tests: exact float32 rounding and start-point-sensitive cases."""

from __future__ import annotations

from fractions import Fraction

import numpy as np

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
