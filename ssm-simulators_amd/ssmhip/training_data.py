"""LAN training-data generation on the device: the TrainingDataGenerator drop-in.

Follows ssms/dataset_generators/lan_mlp.py (TrainingDataGenerator, the
generate_data_training / _save_training_data contract) and the nested
generator configuration of ssms/config/generator_config/data_generator_config.py
(get_lan_kde_config). Where the reference maps one parameter set per worker
process (SimulationPipeline.generate_for_parameter_set), all
n_parameter_sets go through one batched device pass: θ sampling, simulation,
the optional simulation filter with θ resampling, per-choice KDEs, the
resample mixture and the auxiliary labels (ssmhip.lan.LanKDELabeller).
The output dictionary has the reference's keys and float32 arrays, plus the two
configurations; `save=True` writes it as training_data_<uuid>.pickle.
"""

from __future__ import annotations

import logging
from copy import deepcopy

import numpy as np

from . import config as model_configs
from .batched import MODELS
from .lan import LanKDELabeller, sample_theta, save_training_data

logger = logging.getLogger(__name__)

NEG_RT_LABEL = -66.77497


def get_kde_simulation_filters() -> dict:
    """The reference's suggested filter thresholds (data_generator_config.py:7-15)."""
    return {"mode": 20, "choice_cnt": 0, "mean_rt": 17, "std": 0, "mode_cnt_rel": 0.95}


def get_lan_kde_config(model: str = "ddm") -> dict:
    """Nested LAN-KDE generator configuration with the reference's sections and defaults
    (data_generator_config.py:18-75)."""
    return {
        "pipeline": {"n_parameter_sets": 1_000, "n_parameter_sets_rejected": 100, "n_subruns": 10, "n_cpus": "all",
                     "simulation_filters": get_kde_simulation_filters()},
        "estimator": {"type": "kde", "kde_displace_t": False, "pdf_interpolation": "cubic",
                      "max_undecided_prob": 0.5},
        "training": {"n_samples_per_param": 1_000, "mixture_probabilities": [0.8, 0.1, 0.1],
                     "separate_response_channels": False, "negative_rt_log_likelihood": NEG_RT_LABEL},
        "simulator": {"n_samples": 1_000, "delta_t": 0.001, "max_t": 20.0, "smooth_unif": True},
        "output": {"folder": "data/lan_mlp/", "nbins": 0, "pickle_protocol": 4, "bin_pointwise": False},
        "model": model,
    }


def mixture_counts(n: int, probabilities) -> tuple[int, int, int]:
    """Rows per set from each mixture component: int(n p) each, the rounding remainder to the
    KDE samples (resample_mixture_strategy.py:107-124)."""
    p = list(probabilities) + [0.0] * (3 - len(probabilities))
    n_unif = int(n * p[1])
    n_neg = int(n * p[2])
    n_kde = n - n_unif - n_neg
    if n_kde < 0:
        raise ValueError(f"Rounding error too large: cannot adjust n_kde to {n_kde}. n={n}, p={p}")
    return n_kde, n_unif, n_neg


class TrainingDataGenerator:
    """Drop-in for TrainingDataGenerator(config=generator_config, model_config=...) of
    lan_mlp.py:164-396 on the KDE path; the batch runs on the current GPU."""

    def __init__(self, config: dict, model_config: dict | None = None, seed: int | None = None):
        if not isinstance(config, dict):
            raise TypeError("config must be a nested generator configuration dict")
        self.generator_config = deepcopy(config)
        name = self.generator_config.get("model") or (model_config or {}).get("name")
        if model_config is None:
            if name is None:
                raise ValueError("model_config is required when the generator config names no model")
            model_config = model_configs.get(name)
        self.model_config = deepcopy(model_config)
        est = self.generator_config.setdefault("estimator", {})
        est.setdefault("displace_t", False)
        if est.get("type", "kde") != "kde":
            raise NotImplementedError(f"estimator type {est.get('type')!r}: only 'kde' runs on the device")
        base = self.model_config["name"].replace("_deadline", "")
        if base not in MODELS:
            raise NotImplementedError(f"model {base!r} has no batched device path (available: {sorted(MODELS)})")
        if list(self.model_config["params"]) != list(MODELS[base][2]) and "_no_bias" not in base:
            raise ValueError(f"model_config params {self.model_config['params']} do not match {MODELS[base][2]}")
        self.base = base
        self.seed = int(np.random.default_rng().integers(0, 2**31)) if seed is None else int(seed)
        self._calls = 0

    def _sample_theta(self, n_sets: int, seed: int) -> np.ndarray:
        lo, hi = self.model_config["param_bounds"]
        return sample_theta(n_sets, lo, hi, seed=seed)

    def generate_data_training(self, save: bool = False, verbose: bool = False) -> dict:
        """All parameter sets in one device batch (lan_mlp.py:424-561); returns the
        concatenated float32 arrays and the two configurations."""
        gc = self.generator_config
        n_sets = int(gc["pipeline"]["n_parameter_sets"])
        sim = gc.get("simulator", {})
        tr = gc.get("training", {})
        n_kde, n_unif, n_neg = mixture_counts(int(tr.get("n_samples_per_param", 1000)),
                                              tr.get("mixture_probabilities", [0.8, 0.1, 0.1]))
        seed = self.seed + 7919 * self._calls
        self._calls += 1
        theta = self._sample_theta(n_sets, seed)
        lab = LanKDELabeller(self.base, theta, n_samples=int(sim.get("n_samples", 1000)), n_kde=n_kde,
                             n_unif=n_unif, n_neg=n_neg, delta_t=float(sim.get("delta_t", 0.001)),
                             max_t=float(sim.get("max_t", 20.0)), seed=seed + 1,
                             smooth_unif=bool(sim.get("smooth_unif", True)),
                             filters=sim.get("filters"), theta_sampler=self._sample_theta)
        rows = lab.run()
        data = lab.output_dict(rows)
        if verbose and lab.n_resampled:
            logger.info("resampled %d parameter sets rejected by the simulation filter", lab.n_resampled)
        data["generator_config"] = gc
        data["model_config"] = self.model_config
        if save:
            out = gc.get("output", {})
            path = save_training_data(data, out.get("folder", "data/lan_mlp/"), out.get("pickle_protocol", 4))
            if verbose:
                logger.info("saved %s", path)
        return data
