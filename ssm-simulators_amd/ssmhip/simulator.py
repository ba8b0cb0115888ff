"""``Simulator(model).simulate(theta, n_samples, ...)`` — the reference's
class-based entry point (ssms/basic_simulators/simulator_class.py:77-683) with
the GPU simulators of ``ssmhip.cssm`` behind it, plus the legacy functional
``simulator(theta, model, ...)`` (ssms/basic_simulators/simulator.py:624).

Theta preprocessing, deadline handling, noise vectors, parameter transforms,
boundary dictionaries, the simulator call and the output squeeze follow the
reference step by step (simulator.py:80-339, 583-621).
"""

from __future__ import annotations

import numbers
import warnings
from copy import deepcopy

import numpy as np

from . import boundary_functions as bf
from . import config as cfgmod

OMISSION_SENTINEL = -999.0
_C_LONG_MIN, _C_LONG_MAX = -(2**31), 2**31 - 1
_seed_rng = np.random.default_rng()


def _unique_seed() -> int:
    return int(_seed_rng.integers(0, 2**31))


def _validate_random_state(random_state) -> None:
    # simulator.py:58-77
    if random_state is None or not isinstance(random_state, numbers.Integral):
        return
    rs = int(random_state)
    if rs < _C_LONG_MIN or rs > _C_LONG_MAX:
        raise ValueError(f"random_state={rs} is outside [{_C_LONG_MIN}, {_C_LONG_MAX}], the range the C-level "
                         "simulator can use on all platforms.")


def _make_valid_dict(d: dict) -> dict:
    # simulator.py:80-131
    lengths = []
    for key, value in d.items():
        if isinstance(value, list):
            d[key] = np.array(value).astype(np.float32)
        elif isinstance(value, (int, float)):
            d[key] = np.array([value]).astype(np.float32)
        d[key] = np.squeeze(d[key]).astype(np.float32)
        if d[key].ndim > 1:
            raise ValueError(f"Dimension of {key} is greater than 1")
        elif d[key].ndim > 0:
            lengths.append(d[key].shape[0])
    if len(set(lengths)) > 1:
        raise ValueError("thetas have to be either scalars or same length for all thetas which are not scalars")
    if lengths:
        for key, value in d.items():
            if value.ndim == 0:
                d[key] = np.tile(value, lengths[0])
    else:
        for key, value in d.items():
            if value.ndim == 0:
                d[key] = np.expand_dims(value, axis=0)
    return d


class OwnedTheta(np.ndarray):
    """A theta matrix handed over by its owner (hssm_support.rng_fn builds one per call):
    Simulator.simulate uses a float32 one as it is instead of taking the reference's
    defensive float32 copy, since no other code holds it."""


def _preprocess_theta(theta):
    # simulator.py:221-254
    try:
        import pandas as pd
    except Exception:  # pragma: no cover
        pd = None
    if isinstance(theta, list):
        return np.asarray(theta).astype(np.float32)
    if isinstance(theta, np.ndarray):
        if type(theta) is OwnedTheta and theta.dtype == np.float32:
            return theta.view(np.ndarray)
        return theta.astype(np.float32)
    if isinstance(theta, dict):
        return _make_valid_dict(deepcopy(theta))
    if pd is not None and isinstance(theta, pd.DataFrame):
        return {k: np.array(v).astype(np.float32) for k, v in theta.to_dict(orient="list").items()}
    raise ValueError("theta is not supplied as list, numpy array, dictionary, or pandas DataFrame!")


def _theta_array_to_dict(theta: np.ndarray, params: list) -> dict:
    # simulator.py:177-218
    if theta.ndim == 1 and len(params) != theta.shape[0]:
        raise ValueError("model_param_list and theta array do not imply the same number of parameters")
    if theta.ndim == 2 and len(params) != theta.shape[1]:
        raise ValueError("model_param_list and theta array do not imply the same number of parameters")
    if theta.ndim == 1:
        theta = np.expand_dims(theta, axis=0).astype(np.float32)
    return {p: theta[:, i] for i, p in enumerate(params)}


def make_noise_vec(sigma_noise, n_trials: int, n_particles: int) -> np.ndarray:
    # simulator.py:583-621
    shape = n_trials if (n_particles == 1 or n_particles is None) else (n_trials, 1)
    base = [sigma_noise[0]] * n_particles if isinstance(sigma_noise, np.ndarray) else [sigma_noise] * n_particles
    return np.tile(np.array(base, dtype=np.float32), shape)


def make_boundary_dict(config: dict, theta: dict) -> dict:
    # simulator.py:285-339
    fun = config.get("boundary", bf.constant)
    names = config.get("boundary_params") or bf.BOUNDARY_REGISTRY.get(config.get("boundary_name", "constant"),
                                                                       (None, ["a"]))[1]
    if config.get("boundary_name") in bf.BOUNDARY_REGISTRY and not callable(config.get("boundary")):
        fun = bf.BOUNDARY_REGISTRY[config["boundary_name"]][0]
    params = {k: v for k, v in theta.items() if k in names}
    return {"boundary_params": params, "boundary_fun": fun}


class Simulator:
    """Class-based simulator with GPU-backed models (simulator_class.py:34-721)."""

    def __init__(self, model=None, boundary=None, drift=None, simulator_function=None, parameter_adapter=None,
                 parameter_adaptations=None, **config_overrides):
        if drift is not None:
            raise NotImplementedError("time-varying drift functions are not on the GPU path")
        self._adapter = parameter_adapter
        self._custom_adaptations = parameter_adaptations or []
        self._config = self._build_config(model, boundary, simulator_function, config_overrides)

    def _build_config(self, model, boundary, simulator_function, overrides) -> dict:
        if isinstance(model, dict):
            config = deepcopy(model)
            config.setdefault("simulation_transforms", [])
        elif isinstance(model, str):
            config = cfgmod.get(model)
        elif model is None and simulator_function is not None:
            if "params" not in overrides:
                raise ValueError("When using a custom simulator_function without a base model, you must provide "
                                 "'params' (list of parameter names)")
            if "nchoices" not in overrides:
                raise ValueError("When using a custom simulator_function without a base model, you must provide "
                                 "'nchoices' (number of choices)")
            config = {"name": overrides.get("name", "custom"), "params": overrides["params"],
                      "n_params": len(overrides["params"]), "nchoices": overrides["nchoices"],
                      "choices": overrides.get("choices", list(range(overrides["nchoices"]))),
                      "n_particles": overrides.get("n_particles", 1), "simulation_transforms": []}
        else:
            raise ValueError("Must provide either a model name, config dict, or simulator_function")
        config.update(overrides)
        if simulator_function is not None:
            self._validate_simulator_function(simulator_function, config)
            config["simulator"] = simulator_function
        if boundary is not None:
            self._apply_custom_boundary(config, boundary)
        for field in ("params", "nchoices", "simulator"):
            if field not in config:
                raise ValueError(f"Configuration missing required field: '{field}'")
        return config

    @staticmethod
    def _validate_simulator_function(func, config):
        import inspect

        params = inspect.signature(func).parameters
        for p in ("max_t", "n_samples", "delta_t", "random_state"):
            if p not in params:
                warnings.warn(f"Simulator function '{func.__name__}' is missing recommended parameter '{p}'. "
                              "This may cause issues during simulation.", UserWarning)
        for p in config.get("params", []):
            if p not in params:
                raise ValueError(f"Simulator function '{func.__name__}' is missing required model parameter '{p}'")

    @staticmethod
    def _apply_custom_boundary(config, boundary):
        # simulator_class.py:360-402
        if isinstance(boundary, str):
            if boundary not in bf.BOUNDARY_REGISTRY:
                raise ValueError(f"Unknown boundary '{boundary}'. Available boundaries: {list(bf.BOUNDARY_REGISTRY)}")
            fun, names, _ = bf.BOUNDARY_REGISTRY[boundary]
            config["boundary_name"] = boundary
            config["boundary"] = fun
            config["boundary_params"] = list(names)
        elif callable(boundary):
            import inspect

            ps = list(inspect.signature(boundary).parameters)
            if not ps or ps[0] != "t":
                raise ValueError(f"Boundary function '{boundary.__name__}' must accept 't' as first positional argument")
            config["boundary"] = boundary
            config["boundary_name"] = getattr(boundary, "__name__", "custom")
            if "boundary_params" not in config or not config["boundary_params"]:
                if "boundary_params" not in config:
                    warnings.warn("Custom boundary function provided without 'boundary_params'. You may need to "
                                  "specify boundary parameter names.", UserWarning)
                config["boundary_params"] = config.get("boundary_params", [])
        else:
            raise ValueError("boundary must be a string name or callable function")

    @property
    def config(self) -> dict:
        return deepcopy(self._config)

    def simulate(self, theta, n_samples: int = 1000, delta_t: float = 0.001, max_t: float = 20, no_noise: bool = False,
                 sigma_noise=None, smooth_unif: bool = True, random_state=None, return_option: str = "full",
                 n_threads: int = 1, extra_fields=None) -> dict:
        """simulator_class.py:511-683."""
        cfg = deepcopy(self._config)
        model_name = cfg.get("name", "").replace("_deadline", "")
        is_deadline = "deadline" in cfg.get("params", [])
        if random_state is None:
            random_state = _unique_seed()
        _validate_random_state(random_state)
        theta = _preprocess_theta(theta)
        if not isinstance(theta, dict):
            theta = _theta_array_to_dict(theta, cfg["params"])
        n_trials = theta[cfg["params"][0]].shape[0]
        if not is_deadline:
            theta["deadline"] = np.tile(np.array([999], dtype=np.float32), n_trials)
        sim_params = {"max_t": max_t, "n_samples": n_samples, "n_trials": n_trials, "delta_t": delta_t,
                      "random_state": random_state, "return_option": return_option, "smooth_unif": smooth_unif,
                      "n_threads": n_threads}
        if "sd" in theta or "s" in theta:
            if sigma_noise is not None:
                raise ValueError("sigma_noise parameter should be None if 'sd' or 's' is passed via theta dictionary")
            elif no_noise:
                sigma_noise = 0.0
            elif "sd" in theta:
                sigma_noise = theta["sd"]
            else:
                sigma_noise = theta["s"]
        else:
            if no_noise:
                sigma_noise = 0.0
            elif "lba" in model_name and sigma_noise is None:
                sigma_noise = 0.1
            elif sigma_noise is None:
                sigma_noise = 1.0
        noise = make_noise_vec(sigma_noise, n_trials, cfg.get("n_particles", 1))
        theta["sd" if "lba" in model_name else "s"] = noise
        if self._adapter is not None:
            theta = self._adapter.adapt_parameters(theta, cfg, n_trials)
        else:
            for tr in cfg.get("simulation_transforms", []):
                theta = tr(theta, cfg, n_trials)
        for ad in self._custom_adaptations:
            theta = ad.apply(theta, cfg, n_trials) if hasattr(ad, "apply") else ad(theta, cfg, n_trials)
        boundary_dict = make_boundary_dict(cfg, theta) if ("boundary_name" in cfg or "boundary" in cfg) else {}
        if model_name in ("lba2", "lba3", "lba4", "lba_angle_3"):
            if np.any(theta["z"] >= theta["a"]):
                raise ValueError("Starting point z >= a for at least one trial")
        x = cfg["simulator"](**theta, **boundary_dict, **sim_params, **(extra_fields or {}))
        if not isinstance(x, dict):
            raise TypeError(f"Simulator returned {type(x).__name__}, expected dict")
        if n_trials == 1:
            x["rts"] = np.squeeze(x["rts"], axis=1)
            x["choices"] = np.squeeze(x["choices"], axis=1)
        if n_trials > 1 and n_samples == 1:
            x["rts"] = np.squeeze(x["rts"], axis=0)
            x["choices"] = np.squeeze(x["choices"], axis=0)
        x["metadata"]["model"] = model_name
        return x


def simulator(theta, model: str = "angle", n_samples: int = 1000, delta_t: float = 0.001, max_t: float = 20,
              no_noise: bool = False, sigma_noise=None, smooth_unif: bool = True, random_state=None,
              return_option: str = "full", n_threads: int = 1, extra_fields=None) -> dict:
    """Functional entry point (simulator.py:624); returns the class path's core output."""
    return Simulator(model).simulate(theta, n_samples=n_samples, delta_t=delta_t, max_t=max_t, no_noise=no_noise,
                                     sigma_noise=sigma_noise, smooth_unif=smooth_unif, random_state=random_state,
                                     return_option=return_option, n_threads=n_threads, extra_fields=extra_fields)
