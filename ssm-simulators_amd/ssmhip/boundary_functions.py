"""Boundary functions b(t) with the reference's float32 NumPy semantics.

Mirrors ssms/basic_simulators/boundary_functions.py:14-162. The simulators
evaluate `constant`, `angle` and `weibull_cdf` in registers on the GPU; any
other callable with signature ``f(t, **params)`` is evaluated here on the host
over ``t_s`` and handed to the kernel as a per-set table.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def constant(t=0, a: float = 1.0):
    """b(t) = a (boundary_functions.py:14-29)."""
    if isinstance(t, np.ndarray):
        return np.full_like(t, a)
    return a


def angle(t=1, a: float = 1.0, theta: float = 1.0):
    """b(t) = a + t * (-sin(theta)/cos(theta)), i.e. a - t tan(theta) (boundary_functions.py:57-72)."""
    return a + np.multiply(t, (-np.sin(theta) / np.cos(theta)))


def weibull_cdf(t=1, a: float = 1.0, alpha: float = 1.0, beta: float = 1.0):
    """b(t) = a * exp(-(t/beta)^alpha) (boundary_functions.py:103-122)."""
    return a * np.exp(-np.power(np.divide(t, beta), alpha))


def addm_collapse(t=0, a: float = 1.0, b: float = 0.5):
    """b(t) = a - b t, the attentional DDM's collapsing bound (boundary_functions.py:33-54)."""
    return a - np.multiply(t, b)


def generalized_logistic(t=1, a: float = 1.0, B: float = 2.0, M: float = 3.0, v: float = 0.5):  # noqa: N803
    """b(t) = a + 1 - 1 / (1 + exp(-B (t - M)))^(1/v) (boundary_functions.py:76-100)."""
    offset = 1 - (1 / np.power(1 + np.exp(-B * (t - M)), 1 / v))
    return a + offset


def conflict_gamma(t=None, a: float = 1.0, theta: float = 0.5, scale: float = 1.0, alphaGamma: float = 1.01,  # noqa: N803
                   scaleGamma: float = 0.3):  # noqa: N803
    """b(t) = a + scale * GammaPDF(t; alphaGamma, scaleGamma) + t * (-sin(theta)/cos(theta))
    (boundary_functions.py:125-162); t defaults to arange(0, 20, 0.1) as there."""
    from scipy.stats import gamma  # noqa: PLC0415

    if t is None:
        t = np.arange(0, 20, 0.1)
    gamma_bump = scale * gamma.pdf(t, a=alphaGamma, loc=0, scale=scaleGamma)
    angle_offset = np.multiply(t, (-np.sin(theta) / np.cos(theta)))
    return a + gamma_bump + angle_offset


# name -> (function, parameter names, in-kernel boundary code); the last three are evaluated
# on the host into a per-set table
BOUNDARY_REGISTRY = {
    "constant": (constant, ["a"], _lib.BOUNDARY_CONSTANT),
    "angle": (angle, ["a", "theta"], _lib.BOUNDARY_ANGLE),
    "weibull_cdf": (weibull_cdf, ["a", "alpha", "beta"], _lib.BOUNDARY_WEIBULL),
    "addm_collapse": (addm_collapse, ["a", "b"], _lib.BOUNDARY_TABLE),
    "generalized_logistic": (generalized_logistic, ["a", "B", "M", "v"], _lib.BOUNDARY_TABLE),
    "conflict_gamma": (conflict_gamma, ["a", "theta", "scale", "alphaGamma", "scaleGamma"], _lib.BOUNDARY_TABLE),
}

_REFERENCE_MODULE = "ssms.basic_simulators.boundary_functions"


def kernel_boundary_code(fun) -> int:
    """In-register boundary code for `fun`, or BOUNDARY_TABLE for anything else.

    Only this module's functions and the reference module's functions of the
    same name are evaluated in registers; every other callable (including a
    user function that happens to be called "angle") goes through a table.
    """
    if fun is None:
        return _lib.BOUNDARY_CONSTANT
    name = getattr(fun, "__name__", None)
    mod = getattr(fun, "__module__", None)
    if name in BOUNDARY_REGISTRY and (fun is BOUNDARY_REGISTRY[name][0] or mod == _REFERENCE_MODULE):
        return BOUNDARY_REGISTRY[name][2]
    return _lib.BOUNDARY_TABLE


def time_grid(max_t: float, delta_t: float) -> np.ndarray:
    """t_s exactly as _utils.pyx:636-640 builds it (float32 of a float64 arange)."""
    max_t32 = np.float32(max_t)
    dt32 = np.float32(delta_t)
    stop = float(np.float32(max_t32 + dt32))
    return np.arange(0, stop, float(dt32)).astype(np.float32)


def n_steps(max_t: float, delta_t: float) -> int:
    """Length of t_s, i.e. of the boundary array the sequential loop indexes (ddm_models.pyx:668-669).

    The walk index never reaches it at the default max_t / delta_t (t exceeds
    max_t first); the kernel stops a walk there instead of reading past it.
    """
    return int(time_grid(max_t, delta_t).shape[0])
