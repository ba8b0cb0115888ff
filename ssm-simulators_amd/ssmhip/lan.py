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

import ctypes
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
    if kind in (L.MODEL_PAR2, L.MODEL_SEQ2):
        return [0, 1, 2, 3]
    return [-1, 1] if kind in (L.MODEL_DDM, L.MODEL_FULL_DDM, L.MODEL_ORNSTEIN, L.MODEL_LEVY) else list(range(P))


class LanKDELabeller:
    """KDE-labelled LAN rows for a batch of parameter sets, entirely on one GPU."""

    def __init__(self, model: str, theta: np.ndarray, *, n_samples: int = 100_000, n_kde: int = 800,
                 n_unif: int = 100, n_neg: int = 100, delta_t: float = 0.001, max_t: float = 20.0, seed: int = 0,
                 smooth_unif: bool = True, shuffle: bool = True, launch="auto", lb: float = -66.774,
                 tune_sample: float | None = 8e6, filters: dict | None = None, theta_sampler=None,
                 max_attempts: int = 1000):
        """`filters`: the simulation filter thresholds (generator_config["simulator"]["filters"]);
        a set that fails them is re-drawn with `theta_sampler(n, seed)` and re-simulated, up
        to `max_attempts` times (SimulationPipeline.generate_for_parameter_set,
        simulation_pipeline.py:103-115); sets that never pass are dropped from the output."""
        if filters and theta_sampler is None:
            raise ValueError("a simulation filter needs a theta_sampler to re-draw rejected sets")
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
        self.filters = dict(filters) if filters else None
        self.theta_sampler = theta_sampler
        self.max_attempts = int(max_attempts)
        self.n_samples = n_samples
        self.sim_kwargs = dict(delta_t=delta_t, max_t=max_t, smooth_unif=smooth_unif, choice_dtype="int8")
        self.n_resampled = 0
        self.kept = np.ones(self.n_sets, bool)

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
            if self.filters:
                self._apply_filter()
        mark(1)
        bank = KDEBank(self.sim.rts, self.sim.choices, self.sim.job.offsets, self.choice_values,
                       offsets_host=self.sim.offsets)
        mark(2)
        pts = bank.sample(self.n_kde, self.n_unif, max_t=self.max_t, unif_lo=1e-4, seed=self.seed + 1)
        mark(3)
        ll = bank.evaluate(pts, max_t=self.max_t, lb=self.lb)
        mark(4)
        m = pts.m
        W = self.theta.shape[1] + 3
        M = m + self.n_neg  # rows per set: the evaluated points, then the negative-RT rows
        rows = D.empty((self.n_sets * M, W), t.float32)
        # one kernel: the evaluated points, the negative-RT rows (resample_mixture_strategy.py:
        # 202-209) and the shuffle of all of a set's rows (:211-212)
        rc = L.load().ssm_lan_assemble(self.n_sets, self.theta.shape[1], m, self.n_neg, len(self.choice_values),
                                       NEG_RT_LL, int(self.shuffle), self.seed + 2, D.ptr(self.d_theta),
                                       D.ptr(pts.rt), D.ptr(pts.choice), D.ptr(ll), D.ptr(self.d_choice_values),
                                       D.ptr(rows), D.stream_ptr())
        L.check(rc, "ssm_lan_assemble")
        mark(5)
        self.bank, self.points, self.ll = bank, pts, ll
        if not self.kept.all():  # sets the filter rejected on every attempt carry no rows
            keep = np.flatnonzero(self.kept).astype(np.int64)
            out = D.empty((keep.size * M, W), t.float32)
            rc = L.load().ssm_copy_rows(int(keep.size), M * W * 4, D.ptr(rows), D.ptr(D.to_dev(keep)), D.ptr(out),
                                        None, D.stream_ptr())
            L.check(rc, "ssm_copy_rows")
            rows = out
        if ev is not None:
            ev[5].synchronize()
            self.times = StageTimes(ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3]),
                                    ev[3].elapsed_time(ev[4]), ev[4].elapsed_time(ev[5]))
        return rows

    def filter_stats(self, sim=None) -> np.ndarray:
        """Per (set, choice) (mode, mean, std, mode share, count) of a simulation's rts
        without omissions (ssm_filter_stats), float64 [n_sets, n_choices, 5] on the host."""
        t = D.require_gpu()
        sim = sim or self.sim
        S, C = sim.n_sets, len(self.choice_values)
        nb = ctypes.c_size_t(0)
        L.check(L.load().ssm_filter_workspace_bytes(S, sim.n_total, ctypes.byref(nb)), "ssm_filter_workspace_bytes")
        ws = D.empty((max(1, nb.value),), t.uint8)
        stats = D.empty((S, C, 5), t.float64)
        rc = L.load().ssm_filter_stats(S, C, D.ptr(self.d_choice_values), D.ptr(sim.rts), D.ptr(sim.choices),
                                       sim.job.prob.choice_bytes, D.ptr(sim.job.offsets), sim.n_total, D.ptr(ws),
                                       nb.value, D.ptr(stats), D.stream_ptr())
        L.check(rc, "ssm_filter_stats")
        return stats.cpu().numpy()

    def _keep(self, stats: np.ndarray) -> np.ndarray:
        """The keep rule of simulation_pipeline.py:239-246 for every set at once."""
        f = self.filters
        mode_, mean_, std_, rel_, n_ = (stats[:, :, i] for i in range(5))
        ok = ((mode_ <= f.get("mode", np.inf)) & (mean_ <= f.get("mean_rt", np.inf)) & (std_ >= f.get("std", 0.0))
              & (rel_ <= f.get("mode_cnt_rel", 1.0)) & (n_ >= f.get("choice_cnt", 0)))
        return ok.all(axis=1)

    def _apply_filter(self) -> None:
        """Re-draw and re-simulate the rejected sets until they pass or the attempts run out;
        accepted replacements are copied into the batch's rts / choices and theta."""
        t = D.torch()
        n = self.sim.offsets[1:] - self.sim.offsets[:-1]
        if not np.all(n == n[0]):
            raise ValueError("the simulation filter needs the same n_samples for every set")
        n0 = int(n[0])
        rejected = np.flatnonzero(~self._keep(self.filter_stats()))
        rts = self.sim.rts.view(self.n_sets, n0)
        chs = self.sim.choices.view(self.n_sets, n0)
        attempt = 0
        while rejected.size and attempt < self.max_attempts:
            attempt += 1
            th = np.ascontiguousarray(self.theta_sampler(rejected.size, self.seed + 104_729 * attempt), np.float32)
            sub = BatchedSimulation(self.model, th, n0, seed=self.seed + 7 * attempt, launch=self.sim.launch_params,
                                    **self.sim_kwargs)
            sub.run()
            ok = self._keep(self.filter_stats(sub))
            if ok.any():
                src = D.to_dev(np.flatnonzero(ok).astype(np.int64))
                dst = D.to_dev(rejected[ok].astype(np.int64))
                lib = L.load()
                for a, b, nbytes in ((sub.rts, rts, 4), (sub.choices, chs, sub.choices.element_size())):
                    L.check(lib.ssm_copy_rows(int(ok.sum()), n0 * nbytes, D.ptr(a), D.ptr(src), D.ptr(b),
                                              D.ptr(dst), D.stream_ptr()), "ssm_copy_rows")
                self.theta[rejected[ok]] = th[ok]
                self.n_resampled += int(ok.sum())
            sub.close()
            rejected = rejected[~ok]
        self.kept[:] = True
        self.kept[rejected] = False
        self.d_theta = D.to_dev(self.theta)
        t.cuda.current_stream().synchronize()

    def aux_counts(self) -> dict:
        """Integer counts behind the auxiliary labels of every set, from the last simulation
        (SimulationPipeline._compute_auxiliary_labels, simulation_pipeline.py:258-351):
        per-choice draws [S, C], no-go draws [S], RT histograms [S, 128, C] and [S, 256, C]."""
        t = D.require_gpu()
        S, C = self.n_sets, len(self.choice_values)
        counts = D.empty((S, C), t.int32)
        nogo = D.empty((S,), t.int32)
        h128 = D.empty((S, 128, C), t.int32)
        h256 = D.empty((S, 256, C), t.int32)
        rc = L.load().ssm_aux_labels(S, C, D.ptr(self.d_choice_values), int(max(self.choice_values)),
                                     D.ptr(self.sim.rts), D.ptr(self.sim.choices), self.sim.job.prob.choice_bytes,
                                     D.ptr(self.sim.job.offsets), float(self.max_t), D.ptr(counts), D.ptr(nogo),
                                     D.ptr(h128), D.ptr(h256), D.stream_ptr())
        L.check(rc, "ssm_aux_labels")
        return {"counts": counts, "nogo": nogo, "binned_128": h128, "binned_256": h256}

    def aux_labels(self) -> dict:
        """The auxiliary label arrays of all sets, concatenated as TrainingDataGenerator does
        (float32; lan_mlp.py:533-541): cpn, cpn_no_omission, opn, gonogo labels and the
        binned histograms. Probabilities are formed in float64 from exact counts, as the
        reference's np.sum(...) / n_samples, then cast."""
        a = {k: v.cpu().numpy() for k, v in self.aux_counts().items()}
        n = np.asarray(self.sim.offsets[1:] - self.sim.offsets[:-1], np.int64)
        nf = np.maximum(n, 1).astype(np.float64)[:, None]
        choice_p = a["counts"].astype(np.float64) / nf
        # choices never take the omission sentinel (only rts do), so no draw is removed
        choice_p_no_omission = choice_p.copy()
        omission_p = np.zeros((self.n_sets, 1))
        nogo_p = a["nogo"].astype(np.float64)[:, None] / nf
        if len(self.choice_values) == 2:
            cpn = choice_p[:, 1:2]
            cpn_no = choice_p_no_omission[:, 1:2]
        else:
            cpn, cpn_no = choice_p, choice_p_no_omission
        th = self.theta.astype(np.float32)
        f32 = lambda x: np.asarray(x).astype(np.float32)  # noqa: E731
        return {"cpn_data": th, "cpn_labels": f32(cpn), "cpn_no_omission_data": th,
                "cpn_no_omission_labels": f32(cpn_no), "opn_data": th, "opn_labels": f32(omission_p),
                "gonogo_data": th, "gonogo_labels": f32(nogo_p), "binned_128": f32(a["binned_128"]),
                "binned_256": f32(a["binned_256"]), "theta": th}

    def output_dict(self, rows, aux: bool = True) -> dict:
        """The per-key arrays TrainingDataGenerator returns for all sets (lan_mlp.py:533-549):
        lan_data / lan_labels from `rows`, and with `aux` the auxiliary labels of the same
        simulations (call before the next simulate)."""
        r = rows.cpu().numpy()
        out = {"lan_data": r[:, :-1], "lan_labels": r[:, -1], "theta": self.theta[self.kept].astype(np.float32)}
        if aux:
            out.update({k: v[self.kept] for k, v in self.aux_labels().items()})
        return out


def save_training_data(data: dict, output_folder, pickle_protocol: int = 4, generator_config: dict | None = None,
                       model_config: dict | None = None):
    """Write a training-data dict to `output_folder/training_data_<uuid1 hex>.pickle` as
    TrainingDataGenerator._save_training_data does (lan_mlp.py:375-396); with the configs
    given, they are stored under "generator_config" / "model_config" as the reference adds
    them (lan_mlp.py:551-557). Returns the path."""
    import pickle
    import uuid
    from pathlib import Path

    folder = Path(output_folder)
    folder.mkdir(parents=True, exist_ok=True)
    out = dict(data)
    if generator_config is not None:
        out["generator_config"] = generator_config
    if model_config is not None:
        out["model_config"] = model_config
    path = folder / f"training_data_{uuid.uuid1().hex}.pickle"
    with path.open("wb") as f:
        pickle.dump(out, f, protocol=int(pickle_protocol))
    return path


def sample_theta(n_sets: int, lower, upper, seed: int = 0) -> np.ndarray:
    """Parameter sets drawn on the device, independent uniform bounds per parameter
    (UniformParameterSampler, parameter_samplers/uniform_sampler.py:24-51; float64 draw,
    float32 result). Returns float32 [n_sets, n_params] on the host."""
    t = D.require_gpu()
    lo = np.asarray(lower, np.float64).reshape(-1)
    hi = np.asarray(upper, np.float64).reshape(-1)
    if lo.shape != hi.shape or lo.size < 1:
        raise ValueError("lower and upper must have the same, non-zero length")
    if np.any(hi < lo):
        raise ValueError("every upper bound must be >= its lower bound")
    out = D.empty((int(n_sets), lo.size), t.float32)
    d_lo, d_hi = D.to_dev(lo), D.to_dev(hi)
    L.check(L.load().ssm_theta_uniform(int(n_sets), lo.size, D.ptr(d_lo), D.ptr(d_hi), int(seed) & (2**64 - 1),
                                       D.ptr(out), D.stream_ptr()), "ssm_theta_uniform")
    return out.cpu().numpy()


def sample_angle_uniform(n_sets: int = 1024, seed: int = 1):
    """1024 angle-DDM parameter sets, uniform within the angle bounds."""
    rng = np.random.default_rng(seed)
    lower = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    upper = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    return rng.uniform(lower, upper, size=(n_sets, 5)).astype(np.float32)
