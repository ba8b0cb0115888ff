"""LAN training rows on the device: simulate -> per-choice KDE -> resample +
uniform points -> KDE log-likelihood -> shuffled rows.

One pass replaces, for a whole batch of parameter sets, what the reference
does per set inside a process pool:
  SimulationPipeline.generate_for_parameter_set
      ssms/dataset_generators/pipelines/simulation_pipeline.py:78-169
  KDEEstimatorBuilder.build -> LogKDE       estimator_builders/kde_builder.py:66-117
  ResampleMixtureStrategy.generate          strategies/resample_mixture_strategy.py:64-214
Rows are [theta..., rt, choice, loglik] float32 (lan_data = rows[:, :-1],
lan_labels = rows[:, -1], simulation_pipeline.py:153-167). The mixture is
n_kde KDE resamples + n_unif uniform rows (rt ~ U(1e-4, max_t), random
choice) + n_neg negative-RT rows labelled -66.77497, 800 / 100 / 100 by default
(the reference's [0.8, 0.1, 0.1] of 1000 rows per set).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _device as D
from . import _lib as L
from .batched import MODELS, BatchedSimulation
from .kde import KDEBank

NEG_RT_LL = -66.77497


@dataclass
class StageTimes:
    simulate_ms: float = 0.0
    fit_ms: float = 0.0
    sample_ms: float = 0.0
    eval_ms: float = 0.0
    assemble_ms: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def total_ms(self) -> float:
        return self.simulate_ms + self.fit_ms + self.sample_ms + self.eval_ms + self.assemble_ms


def choice_values_for(model: str) -> list[int]:
    kind, _, _, P = MODELS[model]
    return [-1, 1] if kind == L.MODEL_DDM else list(range(P))


class LanKDELabeller:
    """KDE-labelled LAN rows for a batch of parameter sets, entirely on one GPU."""

    def __init__(self, model: str, theta: np.ndarray, *, n_samples: int = 100_000, n_kde: int = 800,
                 n_unif: int = 100, n_neg: int = 100, delta_t: float = 0.001, max_t: float = 20.0, seed: int = 0,
                 smooth_unif: bool = True, shuffle: bool = True, launch="auto", lb: float = -66.774,
                 tune_sample: float | None = 8e6):
        self.model = model
        self.theta = np.ascontiguousarray(theta, np.float32)
        self.n_sets = self.theta.shape[0]
        self.n_kde, self.n_unif, self.n_neg = int(n_kde), int(n_unif), int(n_neg)
        self.max_t, self.seed, self.shuffle, self.lb = float(max_t), int(seed), bool(shuffle), float(lb)
        self.choice_values = choice_values_for(model)
        self.sim = BatchedSimulation(model, self.theta, n_samples, delta_t=delta_t, max_t=max_t, seed=seed,
                                     smooth_unif=smooth_unif, choice_dtype="int8", launch=launch,
                                     tune_sample=tune_sample)
        self.d_theta = D.to_dev(self.theta)
        self.d_choice_values = D.to_dev(np.asarray(self.choice_values, np.int32))
        self.times = StageTimes()

    @property
    def rows_per_set(self) -> int:
        return self.n_kde + self.n_unif + self.n_neg

    def run(self, simulate: bool = True, timed: bool = False):
        """Return rows as a device tensor [n_sets * rows_per_set, n_theta + 3]."""
        t = D.require_gpu()
        ev = [t.cuda.Event(enable_timing=True) for _ in range(6)] if timed else None

        def mark(i):
            if ev is not None:
                ev[i].record()

        mark(0)
        if simulate:
            self.sim.run()
        mark(1)
        bank = KDEBank(self.sim.rts, self.sim.choices, self.sim.job.offsets, self.choice_values)
        mark(2)
        pts = bank.sample(self.n_kde, self.n_unif, max_t=self.max_t, unif_lo=1e-4, seed=self.seed + 1)
        mark(3)
        ll = bank.evaluate(pts, max_t=self.max_t, lb=self.lb)
        mark(4)
        m = pts.m
        W = self.theta.shape[1] + 3
        rows = D.empty((self.n_sets * m, W), t.float32)
        shuffle_here = self.shuffle and not self.n_neg  # with negative rows the shuffle follows their append
        rc = L.load().ssm_lan_assemble(self.n_sets, self.theta.shape[1], m, int(shuffle_here), self.seed + 2,
                                       D.ptr(self.d_theta), D.ptr(pts.rt), D.ptr(pts.choice), D.ptr(ll),
                                       D.ptr(self.d_choice_values), D.ptr(rows), D.stream_ptr())
        L.check(rc, "ssm_lan_assemble")
        mark(5)
        self.bank, self.points, self.ll = bank, pts, ll
        if self.n_neg:
            rows = self._append_negative(rows)
        if ev is not None:
            ev[5].synchronize()
            self.times = StageTimes(ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3]),
                                    ev[3].elapsed_time(ev[4]), ev[4].elapsed_time(ev[5]))
        return rows

    def _append_negative(self, rows):
        """Add n_neg negative-RT rows per set: rt ~ U(-1, 1e-4), random choice, label -66.77497
        (resample_mixture_strategy.py:202-209), then shuffle each set's rows
        (resample_mixture_strategy.py:211-212)."""
        t = D.torch()
        gen = t.Generator(device=rows.device)
        gen.manual_seed(self.seed + 3)
        S, P, W = self.n_sets, self.theta.shape[1], rows.shape[1]
        m = rows.shape[0] // S
        neg = t.empty((S, self.n_neg, W), dtype=t.float32, device=rows.device)
        neg[:, :, :P] = self.d_theta[:, None, :]
        neg[:, :, P] = -1.0 + 1.0001 * t.rand((S, self.n_neg), generator=gen, device=rows.device)
        pick = t.randint(0, len(self.choice_values), (S, self.n_neg), generator=gen, device=rows.device)
        neg[:, :, P + 1] = self.d_choice_values[pick].float()
        neg[:, :, P + 2] = NEG_RT_LL
        out = t.cat([rows.view(S, m, W), neg], dim=1)
        if self.shuffle:
            order = t.rand((S, m + self.n_neg), generator=gen, device=rows.device).argsort(dim=1)
            out = t.gather(out, 1, order[:, :, None].expand(-1, -1, W))
        return out.reshape(S * (m + self.n_neg), W).contiguous()

    def output_dict(self, rows) -> dict:
        """The per-key arrays TrainingDataGenerator concatenates (lan_mlp.py:533-549)."""
        r = rows.cpu().numpy()
        return {"lan_data": r[:, :-1], "lan_labels": r[:, -1], "theta": self.theta}


def sample_config4(n_sets: int = 1024, seed: int = 1):
    """1024 angle-DDM parameter sets, uniform within the angle bounds."""
    rng = np.random.default_rng(seed)
    lower = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    upper = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    return rng.uniform(lower, upper, size=(n_sets, 5)).astype(np.float32)
