"""Batched, ragged simulation: many parameter sets, each with its own n_samples,
in ONE persistent kernel launch (optionally captured in a HIP graph).

The reference needs one ``Simulator.simulate`` call per parameter set (its
``n_samples`` is per call, src/cssm/ddm_models.pyx:589-603, and the LAN
generator maps sets over a process pool, ssms/dataset_generators/lan_mlp.py:490-500).
Here the flattened (set, sample) space is addressed through prefix-sum
offsets; outputs are set-major: rts[offsets[k] + n], choices int8 (or int32).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _device as D
from . import _lib as L
from . import autotune
from .boundary_functions import BOUNDARY_REGISTRY
from .em import (LaunchParams, accumulator_records, boundary_table, ddm_records, expected_steps, make_job,
                 two_stage_records)

F32 = np.float32

# model name -> (kernel model, boundary name, parameter order, particles)
MODELS = {
    "ddm": (L.MODEL_DDM, "constant", ["v", "a", "z", "t"], 1),
    "angle": (L.MODEL_DDM, "angle", ["v", "a", "z", "t", "theta"], 1),
    "weibull": (L.MODEL_DDM, "weibull_cdf", ["v", "a", "z", "t", "alpha", "beta"], 1),
    "race_4": (L.MODEL_RACE, "constant", ["v0", "v1", "v2", "v3", "a", "z0", "z1", "z2", "z3", "t"], 4),
    "race_4_weibull": (L.MODEL_RACE, "weibull_cdf",
                       ["v0", "v1", "v2", "v3", "a", "z0", "z1", "z2", "z3", "t", "alpha", "beta"], 4),
    "lca_4": (L.MODEL_LCA, "constant", ["v0", "v1", "v2", "v3", "a", "z0", "z1", "z2", "z3", "g", "b", "t"], 4),
    "full_ddm": (L.MODEL_FULL_DDM, "constant", ["v", "a", "z", "t", "sz", "sv", "st"], 1),
    "ornstein": (L.MODEL_ORNSTEIN, "constant", ["v", "a", "z", "g", "t"], 1),
    "ornstein_angle": (L.MODEL_ORNSTEIN, "angle", ["v", "a", "z", "g", "t", "theta"], 1),
    "levy": (L.MODEL_LEVY, "constant", ["v", "a", "z", "alpha", "t"], 1),
    "levy_angle": (L.MODEL_LEVY, "angle", ["v", "a", "z", "alpha", "t", "theta"], 1),
    "racing_diffusion_3": (L.MODEL_RDM, "constant", ["v0", "v1", "v2", "A", "b", "t"], 3),
    "ddm_par2": (L.MODEL_PAR2, "constant", ["vh", "vl1", "vl2", "a", "zh", "zl1", "zl2", "t"], 1),
    "ddm_seq2": (L.MODEL_SEQ2, "constant", ["vh", "vl1", "vl2", "a", "zh", "zl1", "zl2", "t"], 1),
    "ddm_par2_angle_no_bias": (L.MODEL_PAR2, "angle", ["vh", "vl1", "vl2", "a", "t", "theta"], 1),
    "ddm_seq2_angle_no_bias": (L.MODEL_SEQ2, "angle", ["vh", "vl1", "vl2", "a", "t", "theta"], 1),
}


def records_for(model: str, theta: np.ndarray, delta_t: float, max_t: float, deadline=None, s: float = 1.0):
    """Per-set records from a theta matrix [n_sets, n_params] in the model's parameter order."""
    kind, bname, names, P = MODELS[model]
    theta = np.asarray(theta, F32)
    if theta.ndim != 2 or theta.shape[1] != len(names):
        raise ValueError(f"theta for {model} must have shape (n_sets, {len(names)})")
    n = theta.shape[0]
    col = {nm: theta[:, i] for i, nm in enumerate(names)}
    dl = np.full(n, 999.0, F32) if deadline is None else np.asarray(deadline, F32).reshape(n)
    bnames = BOUNDARY_REGISTRY[bname][1]
    if kind == L.MODEL_RDM:
        col["a"] = col["b"]  # the racing diffusion's threshold b is a constant boundary
        col.update({f"z{j}": col["A"] for j in range(P)})  # start range U(0, A)
    bparams = {k: col[k] for k in bnames}
    code = BOUNDARY_REGISTRY[bname][2]
    if kind in (L.MODEL_PAR2, L.MODEL_SEQ2):
        half = np.full(n, 0.5, F32)  # the no-bias variants fix the starts at 1/2
        zs = [col.get(k, half) for k in ("zh", "zl1", "zl2")]
        rec = two_stage_records(col["vh"], col["vl1"], col["vl2"], *zs, col["t"], dl, np.full(n, s, F32), max_t,
                                delta_t, code, bparams)
    elif kind in (L.MODEL_DDM, L.MODEL_FULL_DDM, L.MODEL_ORNSTEIN, L.MODEL_LEVY):
        extra = {k: col[k] for k in ("sz", "sv", "st", "g") if k in col}
        if kind == L.MODEL_LEVY:
            extra["alpha"] = col["alpha"]
        rec = ddm_records(col["v"], col["a"], col["z"], col["t"], dl, np.full(n, s, F32), max_t, delta_t, code, bparams,
                          **extra)
    else:
        v = np.stack([col[f"v{j}"] for j in range(P)], axis=1)
        z = np.stack([col[f"z{j}"] for j in range(P)], axis=1)
        g = col.get("g")
        b = col.get("b") if kind == L.MODEL_LCA else None
        rec = accumulator_records(v, z, col["t"], np.full((n, 1), s, F32), dl, max_t, delta_t, code, bparams,
                                  g=g, b=b)
    return kind, code, P, rec, bname, bparams


class BatchedSimulation:
    """One persistent launch over all (set, sample) pairs of a ragged batch."""

    def __init__(self, model: str, theta: np.ndarray, n_samples, *, delta_t: float = 0.001, max_t: float = 20.0,
                 seed: int = 0, smooth_unif: bool = False, choice_dtype: str = "int8", launch="auto",
                 deadline=None, replay_normals=None, replay_uniforms=None, use_table: bool = False,
                 tune_sample: float | None = 8e6, queue_order="expected"):
        kind, code, P, rec, bname, bparams = records_for(model, theta, delta_t, max_t, deadline)
        n_sets = rec.shape[0]
        ns = np.broadcast_to(np.asarray(n_samples, np.int64), (n_sets,))
        if np.any(ns < 0):
            raise ValueError("n_samples must be >= 0")
        self.offsets = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        table = None
        if use_table or code == L.BOUNDARY_TABLE:  # boundaries without an in-register form always use one
            table = boundary_table(BOUNDARY_REGISTRY[bname][0], bparams, n_sets, max_t, delta_t)
            code = L.BOUNDARY_TABLE
        self.model = model
        self.n_sets = n_sets
        self.n_total = int(self.offsets[-1])
        self.job = make_job(kind, code, P, rec, self.offsets, delta_t=delta_t, max_t=max_t, seed=seed,
                            smooth_unif=smooth_unif, out_layout=0,
                            choice_bytes=1 if choice_dtype == "int8" else 4, table=table,
                            replay_normals=replay_normals, replay_uniforms=replay_uniforms)
        self._graph = None
        # queue visiting order: "expected" = sets by a mean-walk-length proxy, longest first
        # (shortens the launch tail; outputs are identical in any order), None = index order,
        # or an explicit permutation of the sets
        if isinstance(queue_order, str):
            if queue_order != "expected":
                raise ValueError(f"unknown queue_order {queue_order!r}")
            order = np.argsort(-expected_steps(rec, kind, P, max_t, delta_t), kind="stable")
        else:
            order = queue_order
        self.job.set_queue_order(order)
        if launch == "auto":
            if tune_sample is None:  # tune on the batch itself
                self.launch_params = autotune.tune(self.job)
            else:
                self.launch_params = self._autotune(kind, code, P, rec, delta_t, max_t, seed, smooth_unif, table,
                                                    tune_sample)
        else:
            self.launch_params = launch or LaunchParams()

    def _autotune(self, kind, code, P, rec, delta_t, max_t, seed, smooth_unif, table, tune_sample):
        # tune on a bounded sample: up to 1024 evenly spaced sets (keeps the parameter mix),
        # at most `tune_sample` trajectories
        idx = np.arange(0, self.n_sets, max(1, self.n_sets // 1024))
        ns = np.diff(self.offsets)[idx]
        scale = min(1.0, tune_sample / max(1, ns.sum()))
        ns = np.maximum(1, (ns * scale).astype(np.int64))
        offs = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        sub = make_job(kind, code, P, rec[idx], offs, delta_t=delta_t, max_t=max_t, seed=seed,
                       smooth_unif=smooth_unif, out_layout=0, choice_bytes=1,
                       table=None if table is None else table[idx])
        if self.job.set_order is not None:  # the sample keeps the batch's relative queue order
            rank = np.empty(self.n_sets, np.int64)
            rank[self.job.set_order.cpu().numpy()] = np.arange(self.n_sets)
            sub.set_queue_order(np.argsort(rank[idx], kind="stable"))
        return autotune.tune(sub)

    # ---- eager launch -------------------------------------------------------
    def run(self, stream=None) -> None:
        self.job.launch(self.launch_params, stream)

    # ---- HIP graph ----------------------------------------------------------
    def capture(self) -> None:
        """Capture reset + persistent kernel into a graph executable (pointers baked in)."""
        if self._graph is not None:
            return
        lc = self.job.launch_config(self.launch_params)
        h = C.c_void_p()
        rc = L.load().ssm_em_graph_create(C.byref(self.job.prob), C.byref(lc), *self.job.pointers(), C.byref(h))
        L.check(rc, "ssm_em_graph_create")
        self._graph = h

    def replay(self, stream=None) -> None:
        if self._graph is None:
            self.capture()
        L.check(L.load().ssm_graph_launch(self._graph, D.stream_ptr(stream)), "ssm_graph_launch")

    def close(self) -> None:
        if self._graph is not None:
            L.load().ssm_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    # ---- results ----------------------------------------------------------------
    @property
    def rts(self):
        return self.job.rts[: self.n_total]

    @property
    def choices(self):
        return self.job.choices[: self.n_total]

    def steps(self) -> int:
        return self.job.steps()

    def lane_slots(self) -> int:
        return self.job.lane_slots()

    def per_set(self, k: int):
        lo, hi = int(self.offsets[k]), int(self.offsets[k + 1])
        return self.rts[lo:hi].cpu().numpy(), self.choices[lo:hi].cpu().numpy()


def sample_angle_ragged(n_sets: int = 4096, seed: int = 0, lo: float = 1e3, hi: float = 1e5):
    """Batched angle-DDM workload: theta ~ U(angle bounds), n_i = floor(10^U(log10 lo, log10 hi))."""
    rng = np.random.default_rng(seed)
    lower = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    upper = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    theta = rng.uniform(lower, upper, size=(n_sets, 5)).astype(F32)
    ns = np.floor(10 ** rng.uniform(np.log10(lo), np.log10(hi), n_sets)).astype(np.int64)
    return theta, ns
