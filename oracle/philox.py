"""Philox4x32-10 in uint32/uint64 NumPy and the Box–Muller map (test oracle).

Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11 (Random123). Pinned by the Random123 known-answer vectors
(tests/test_oracle_rng.py). The Box–Muller map is computed in float64 and
compared with the GPU's float32 hardware-transcendental version to a
tolerance, not bit-exactly.
"""

from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key: tuple[int, int], rounds: int = 10) -> np.ndarray:
    """ctr: uint32[..., 4]; returns uint32[..., 4]."""
    c = np.asarray(ctr, dtype=np.uint64)
    c0, c1, c2, c3 = (c[..., i].copy() for i in range(4))
    k0 = np.uint64(key[0]) & _MASK
    k1 = np.uint64(key[1]) & _MASK
    for _ in range(rounds):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def key_from_seed(seed: int) -> tuple[int, int]:
    seed &= 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def open01(u: np.ndarray) -> np.ndarray:
    """(u + 0.5) * 2^-32 in float64: in (0, 1)."""
    return (u.astype(np.float64) + 0.5) * 2.0**-32


def unit(u: np.ndarray) -> np.ndarray:
    """top 24 bits * 2^-24: in [0, 1)."""
    return (u >> np.uint32(8)).astype(np.float64) * 2.0**-24


def rev(u: np.ndarray) -> np.ndarray:
    """Box–Muller angle in revolutions: top 23 bits * 2^-23, in [0, 1). The kernels place the
    same bits under the exponent of 1.0f (a value in [1, 2)) and let the hardware sin/cos take
    the fraction, which is exact."""
    return (u >> np.uint32(9)).astype(np.float64) * 2.0**-23


def box_muller(a: np.ndarray, b: np.ndarray, s: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """Box & Muller (1958): r = sqrt(-2 ln u1) s, angle 2*pi*u2 (float64); s scales both
    normals as the kernels' radius sqrt(-2 ln u1 s^2) does."""
    r = np.sqrt(-2.0 * np.log(open01(a))) * abs(s)
    th = 2.0 * np.pi * rev(b)
    return r * np.cos(th), r * np.sin(th)


def normals4(ctr: np.ndarray, seed: int) -> np.ndarray:
    """Four N(0,1) per counter, in the order the GPU kernels consume them."""
    r = philox4x32_10(ctr, key_from_seed(seed))
    z0, z1 = box_muller(r[..., 0], r[..., 1])
    z2, z3 = box_muller(r[..., 2], r[..., 3])
    return np.stack([z0, z1, z2, z3], axis=-1)
