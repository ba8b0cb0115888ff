"""Closed forms for the constant-bound Wiener first-passage problem (test oracle).

Process dX = v dt + dW on (-a, a), X0 = (2z - 1) a (the ddm_flexbound
convention, src/cssm/ddm_models.pyx:709). Written from the standard results:
  * absorption probability and mean exit time (Cox & Miller 1965, ch. 5),
  * first-passage density series of Navarro & Fuss (2009), with the
    large-time and small-time expansions.
Discrete Euler steps overshoot the boundary; the Siegmund (1979) continuity
correction moves each bound out by beta * sqrt(dt) with beta = -zeta(1/2)/sqrt(2pi)
= 0.5826.
Pinned against scipy-free brute force (numerical integration of the density
equals the absorption probability; tests/test_oracle_em.py).
"""

from __future__ import annotations

import numpy as np

SIEGMUND_BETA = 0.5825971579390106


def p_upper(v: float, a: float, z: float) -> float:
    """P(hit +a before -a) for the ±a bounds and start (2z-1)a."""
    w = 2.0 * a  # separation
    x0 = z * w  # distance from the lower bound
    if abs(v) < 1e-12:
        return x0 / w
    return float((1.0 - np.exp(-2.0 * v * x0)) / (1.0 - np.exp(-2.0 * v * w)))


def mean_decision_time(v: float, a: float, z: float) -> float:
    w = 2.0 * a
    x0 = z * w
    if abs(v) < 1e-12:
        return float(x0 * (w - x0))
    pu = p_upper(v, a, z)
    return float((w * pu - x0) / v)


def _lower_density_std(t: np.ndarray, w: float, kterms: int = 10) -> np.ndarray:
    """Navarro & Fuss density of hitting 0 first for v=0, a=1, start w (t standardised)."""
    t = np.asarray(t, np.float64)
    out = np.zeros_like(t)
    big = t > 0.3
    if np.any(big):
        tb = t[big]
        k = np.arange(1, 60)[:, None]
        out[big] = np.pi * np.sum(k * np.exp(-(k**2) * np.pi**2 * tb / 2.0) * np.sin(k * np.pi * w), axis=0)
    small = (t > 0) & ~big
    if np.any(small):
        ts = t[small]
        k = np.arange(-kterms, kterms + 1)[:, None]
        out[small] = (1.0 / np.sqrt(2 * np.pi * ts**3)) * np.sum(
            (w + 2 * k) * np.exp(-((w + 2 * k) ** 2) / (2.0 * ts)), axis=0
        )
    return out


def density(t: np.ndarray, v: float, a: float, z: float, choice: int) -> np.ndarray:
    """First-passage density at decision time t for choice +1 (upper) or -1 (lower)."""
    w_sep = 2.0 * a
    zr = z if choice < 0 else 1.0 - z
    vv = v if choice < 0 else -v
    t = np.asarray(t, np.float64)
    ts = t / (w_sep**2)
    f = _lower_density_std(ts, zr)
    return f / (w_sep**2) * np.exp(-vv * w_sep * zr - vv * vv * t / 2.0)


def log_density_rt(rt: np.ndarray, choice: np.ndarray, v: float, a: float, z: float, t0: float) -> np.ndarray:
    """log p(rt, choice) with non-decision time t0 (rt <= t0 -> -inf)."""
    rt = np.asarray(rt, np.float64)
    choice = np.asarray(choice)
    out = np.full(rt.shape, -np.inf)
    for c in (-1, 1):
        m = (choice == c) & (rt > t0)
        if np.any(m):
            out[m] = np.log(np.maximum(density(rt[m] - t0, v, a, z, c), 1e-300))
    return out
