"""Posterior/prior-predictive sampling for HSSM on the GPU simulators.

Drop-in for ssms/hssm_support.py: HSSM's random variable calls `rng_fn` with one
parameter array per model parameter (scalars, or arrays whose last axis is the
data trial); the call becomes one `simulator(theta, model, n_samples=n_replicas)`
over all trials, i.e. the extreme ragged case of many parameter sets with one
or a few samples each, which the EM kernel's work queue runs in one launch.
Names, argument meanings, shapes and errors follow the reference module
(hssm_support.py:13-455).
"""

from __future__ import annotations

import logging
from functools import partial
from typing import Any, Callable

import numpy as np

from .config import get as _get_config
from .simulator import OwnedTheta, simulator

_logger = logging.getLogger(__name__)


def _extract_size_val(size: tuple | int) -> int:
    """First entry of a size tuple, or the size itself (hssm_support.py:16-20)."""
    return int(size[0]) if isinstance(size, tuple) else size


def _validate_size(size_val: int, new_data_size: int) -> None:
    """`size` must be a whole number of data sets (hssm_support.py:169-190)."""
    if size_val % new_data_size != 0:
        raise ValueError("`size` needs to be a multiple of the size of data")


def _calculate_n_replicas(is_all_args_scalar, size, new_data_size) -> int:
    """Samples per trial: 1 for scalar parameters or size None/1, else size / n_data_trials
    (hssm_support.py:23-52)."""
    if is_all_args_scalar or size is None or size == 1:
        return 1
    size_val = _extract_size_val(size)
    _validate_size(size_val, new_data_size)
    return size_val // new_data_size


def _get_seed(rng) -> int:
    """A seed in [0, 2**31), the range the simulator's C RNG accepts (hssm_support.py:55-65)."""
    return int(rng.integers(0, 2**31))


def _prepare_theta_and_shape(arg_arrays, size):
    """The (rows, n_params) theta matrix and the output shape (hssm_support.py:68-109).

    All-scalar parameters: one row repeated `size` times (size an int or a 1-tuple).
    Otherwise every parameter is broadcast to the largest one's shape and flattened,
    so row i is trial i % n_data_trials of replica block i // n_data_trials.
    """
    if all(np.asarray(arg).size == 1 for arg in arg_arrays):
        theta = np.stack([np.asarray(arg) for arg in arg_arrays])
        if theta.ndim > 1:
            theta = theta.squeeze(axis=-1)
        if isinstance(size, tuple) and len(size) == 1:
            n_rows = size[0]
        elif isinstance(size, int):
            n_rows = size
        else:
            raise ValueError(f"Size must be a tuple of length 1 or an integer, but got {type(size)}")
        return True, np.tile(theta, (n_rows, 1)), None, None
    arrays = [np.asarray(arg) for arg in arg_arrays]
    max_shape = arrays[int(np.argmax([a.size for a in arrays]))].shape
    # column-major, so that each parameter's column is contiguous for the records built from it
    theta = np.empty((int(np.prod(max_shape)), len(arrays)), dtype=np.result_type(*arrays), order="F")
    for j, arr in enumerate(arrays):
        theta[:, j].reshape(max_shape)[...] = arr
    return False, theta, max_shape, max_shape[-1]


def _broadcast_extra_fields(extra_fields, max_shape, new_data_size) -> dict:
    """Per-trial covariates (leading length n_data_trials) broadcast and flattened like theta;
    others passed through (hssm_support.py:112-136)."""
    lead = tuple(max_shape[:-1])
    out = {}
    for key, value in extra_fields.items():
        arr = np.asarray(value)
        if arr.ndim == 0 or arr.shape[0] != new_data_size:
            out[key] = arr
        else:
            out[key] = np.ascontiguousarray(np.broadcast_to(arr, lead + arr.shape).reshape((-1,) + arr.shape[1:]))
    return out


def _reshape_sims_out(max_shape, n_replicas, obs_dim_int) -> tuple:
    """(*max_shape[, n_replicas], obs_dim) (hssm_support.py:139-166)."""
    shape = [*max_shape]
    if n_replicas != 1:
        shape.append(n_replicas)
    shape.append(obs_dim_int)
    return tuple(shape)


def _validate_simulator_fun_arg(simulator_fun) -> None:
    if not (isinstance(simulator_fun, str) or callable(simulator_fun)):
        raise ValueError(f"`simulator_fun` must be a string or a callable, but you passed {type(simulator_fun)}.")


def decorate_atomic_simulator(model_name: str, choices=None, obs_dim: int = 2):
    """Attach model_name / choices / obs_dim to a simulator callable (hssm_support.py:212-252)."""
    choices = [-1, 1] if choices is None else choices

    def decorator(func):
        func.model_name = model_name
        func.choices = choices
        func.obs_dim = obs_dim
        return func

    return decorator


def hssm_sim_wrapper(simulator_fun, theta, model, n_replicas, random_state, **kwargs):
    """simulator(theta, model, n_samples=n_replicas) -> (rt, response) stacked on the last axis
    and squeezed (hssm_support.py:255-290)."""
    out = simulator_fun(theta=theta, model=model, n_samples=n_replicas, random_state=random_state, **kwargs)
    return np.stack([out["rts"], out["choices"]], axis=-1).squeeze()


def _build_decorated_simulator(model_name: str, choices: list, obs_dim: int = 2) -> Callable:
    wrapped = partial(hssm_sim_wrapper, simulator_fun=simulator, model=model_name)
    wrapped.takes_owned_theta = True  # rng_fn may hand it its per-call theta without a copy
    return decorate_atomic_simulator(model_name=model_name, choices=choices, obs_dim=obs_dim)(wrapped)


def get_simulator_fun_internal(simulator_fun: Callable | str):
    """A callable is returned as given; a model name becomes the decorated GPU simulator
    (hssm_support.py:308-344). Unknown names only warn here, and fail when sampled."""
    _validate_simulator_fun_arg(simulator_fun)
    if callable(simulator_fun):
        return simulator_fun
    try:
        choices = list(_get_config(simulator_fun).get("choices", [0, 1, 2]))
    except ValueError:
        _logger.warning(
            "You supplied a model '%s', which is currently not supported in the ssm_simulators package. An "
            "error will be thrown when sampling from the random variable or when using any posterior or prior "
            "predictive sampling methods.", simulator_fun)
        choices = [0, 1, 2]
    return _build_decorated_simulator(model_name=simulator_fun, choices=choices)


def validate_simulator_fun(simulator_fun: Any) -> tuple[str, list, int]:
    """(model_name, choices, obs_dim) of a decorated simulator (hssm_support.py:347-384)."""
    if not hasattr(simulator_fun, "model_name"):
        raise ValueError("The simulator function must have a `model_name` attribute.")
    if not hasattr(simulator_fun, "choices"):
        raise ValueError("The simulator function must have a `choices` attribute.")
    if not hasattr(simulator_fun, "obs_dim"):
        raise ValueError("The simulator function must have a `obs_dim` attribute.")
    if not isinstance(simulator_fun.obs_dim, int):
        raise ValueError("The obs_dim attribute must be an integer")
    return simulator_fun.model_name, simulator_fun.choices, simulator_fun.obs_dim


def rng_fn(arg_arrays, size, rng, simulator_fun: Callable, obs_dim_int: int, *args, **kwargs) -> np.ndarray:
    """(rt, response) draws for HSSM's random variable (hssm_support.py:388-455): all trials and
    replicas in one simulator call, reshaped to (*param_shape[, n_replicas], obs_dim) for
    vector parameters."""
    is_all_args_scalar, theta, max_shape, new_data_size = _prepare_theta_and_shape(arg_arrays, size)
    n_replicas = _calculate_n_replicas(is_all_args_scalar, size, new_data_size)
    seed = _get_seed(rng)
    if kwargs.get("extra_fields") is not None and not is_all_args_scalar and max_shape is not None:
        kwargs = {**kwargs, "extra_fields": _broadcast_extra_fields(kwargs["extra_fields"], max_shape,
                                                                    new_data_size)}
    if getattr(simulator_fun, "takes_owned_theta", False):
        theta = theta.view(OwnedTheta)  # built above for this call only: no copy needed
    sims_out = simulator_fun(theta=theta, random_state=seed, n_replicas=n_replicas, **kwargs)
    if not is_all_args_scalar:
        sims_out = sims_out.reshape(_reshape_sims_out(max_shape, n_replicas, obs_dim_int))
    return sims_out
