"""Drop-ins for the LAN data-generation protocol objects on the KDE path.

  KDELikelihoodEstimator  ssms/dataset_generators/likelihood_estimators/kde_estimator.py:15-148
  KDEEstimatorBuilder     ssms/dataset_generators/estimator_builders/kde_builder.py:24-117
  ResampleMixtureStrategy ssms/dataset_generators/strategies/resample_mixture_strategy.py:18-214

They satisfy the same protocols (fit / evaluate / sample / get_metadata;
build(theta, simulations); generate(theta, estimator, random_state)) and can be
injected into the reference's SimulationPipeline(estimator_builder=...,
training_strategy=...). The LogKDE behind them fits and evaluates on the GPU.
For whole batches of parameter sets use ssmhip.lan.LanKDELabeller instead.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .kde import LogKDE


class KDELikelihoodEstimator:
    def __init__(self, displace_t: bool = False):
        self.displace_t = displace_t
        self._kde: LogKDE | None = None
        self._metadata: dict[str, Any] | None = None

    def fit(self, simulations: dict[str, Any]) -> None:
        self._kde = LogKDE(simulations, displace_t=self.displace_t)
        self._metadata = simulations["metadata"]

    def evaluate(self, rts: np.ndarray, choices: np.ndarray) -> np.ndarray:
        if self._kde is None:
            raise ValueError("Must call fit() before evaluate()")
        return np.asarray(self._kde.kde_eval(data={"rts": rts, "choices": choices})).ravel()

    def sample(self, n_samples: int, random_state: int | None = None) -> dict[str, np.ndarray]:
        if self._kde is None:
            raise ValueError("Must call fit() before sample()")
        s = self._kde.kde_sample(n_samples=n_samples, random_state=random_state)
        return {"rts": s["rts"].ravel(), "choices": s["choices"].ravel()}

    def get_metadata(self) -> dict[str, Any]:
        if self._metadata is None:
            raise ValueError("Must call fit() before get_metadata()")
        return self._metadata


class KDEEstimatorBuilder:
    def __init__(self, generator_config: dict | None = None):
        cfg = generator_config or {}
        est = cfg.get("estimator", {}) if isinstance(cfg, dict) else {}
        self.displace_t = bool(est.get("displace_t", est.get("kde_displace_t", cfg.get("kde_displace_t", False))))

    def build(self, theta: dict[str, Any], simulations: dict[str, Any] | None = None) -> KDELikelihoodEstimator:
        if simulations is None:
            raise ValueError("KDE estimator requires simulations data. Cannot build KDE estimator with "
                             "simulations=None.")
        est = KDELikelihoodEstimator(displace_t=self.displace_t)
        est.fit(simulations)
        return est


class ResampleMixtureStrategy:
    """[p_kde, p_unif_up, p_unif_down] mixture of resampled and uniform rows (float32, shuffled)."""

    def __init__(self, generator_config: dict, model_config: dict):
        self.generator_config = generator_config
        self.model_config = model_config

    def generate(self, theta: dict[str, Any], likelihood_estimator, random_state: int | None = None) -> np.ndarray:
        rng = np.random.default_rng(random_state)
        tc = self.generator_config.get("training", self.generator_config)
        n = tc.get("n_samples_per_param") or tc.get("n_training_samples_by_parameter_set", 1000)
        p = tc.get("mixture_probabilities") or tc.get("data_mixture_probabilities", [0.8, 0.1, 0.1])
        n_kde, n_up, n_down = int(n * p[0]), int(n * p[1]), int(n * p[2])
        total = n_kde + n_up + n_down
        if total != n:
            n_kde += n - total
            if n_kde < 0:
                raise ValueError(f"Rounding error too large: cannot adjust n_kde to {n_kde}. n={n}, p={p}")
        meta = likelihood_estimator.get_metadata()
        sep = bool(tc.get("separate_response_channels", False))
        nch = self.model_config["nchoices"]
        n_feat = (2 + nch + len(theta)) if sep else (3 + len(theta))
        out = np.zeros((n_kde + n_up + n_down, n_feat))
        vals = [val if np.isscalar(val) else np.asarray(val).flat[0] for val in
                (theta[k] for k in self.model_config["params"])]
        out[:, : len(theta)] = np.tile(np.array([vals]), (out.shape[0], 1))
        s = likelihood_estimator.sample(n_samples=n_kde, random_state=random_state)
        ll = likelihood_estimator.evaluate(s["rts"], s["choices"])
        if sep:
            out[:n_kde, (-2 - nch)] = s["rts"]
            for r_cnt, resp in enumerate(meta["possible_choices"]):
                out[:n_kde, (-1 - nch) + r_cnt] = (s["choices"] == resp).astype(int)
        else:
            out[:n_kde, -3] = s["rts"]
            out[:n_kde, -2] = s["choices"]
        out[:n_kde, -1] = ll
        ch = rng.choice(meta["possible_choices"], size=n_up)
        hi = meta["max_t"] if meta["max_t"] < 100 else 100
        rt = rng.uniform(low=0.0001, high=hi, size=n_up)
        out[n_kde:n_kde + n_up, -3] = rt
        out[n_kde:n_kde + n_up, -2] = ch
        out[n_kde:n_kde + n_up, -1] = likelihood_estimator.evaluate(rt, ch)
        ch = rng.choice(meta["possible_choices"], size=n_down)
        out[n_kde + n_up:, -3] = rng.uniform(low=-1.0, high=0.0001, size=n_down)
        out[n_kde + n_up:, -2] = ch
        out[n_kde + n_up:, -1] = tc.get("negative_rt_log_likelihood", -66.77497)
        rng.shuffle(out)
        return out.astype(np.float32)


MixtureTrainingStrategy = ResampleMixtureStrategy
