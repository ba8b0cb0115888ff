"""Gaussian KDE on log RT, fitted and evaluated on the GPU.

``KDEBank`` holds one KDE per (parameter set, choice) for a whole batch of
simulated sets (the LAN labelling workload); ``LogKDE`` is the drop-in for the
reference class ssms/support_utils/kde_class.py:15-504 (same constructor and
``kde_eval`` / ``kde_sample`` keyword contracts), backed by a one-set bank.
``bandwidth_silverman`` mirrors kde_class.py:508-553 on the host.
"""

from __future__ import annotations

from collections.abc import Iterable
from copy import deepcopy

import numpy as np

from . import _device as D
from . import _lib as L

OMISSION_SENTINEL = -999.0
LB_DEFAULT = -66.774


def bandwidth_silverman(sample: Iterable[float] = (0, 0, 0), std_cutoff: float = 1e-3, std_proc: str = "restrict",
                        std_n_1: float = 10.0) -> np.float64:
    """Silverman bandwidth (4/(3n))^(1/5) * std (kde_class.py:508-553)."""
    s = np.array(sample)
    n = len(s)
    if n == 0:
        raise ValueError("Sample is empty")
    if n > 1:
        std = np.std(s)
        if std < std_cutoff:
            std = std_cutoff if std_proc == "restrict" else 0.0
    else:
        std = std_n_1
    return np.power((4 / (3 * n)), 1 / 5) * std


_CHOICE_VALUES: dict = {}


def _choice_values_dev(values: tuple):
    """The possible choice values as a device int32 array, uploaded once per device."""
    key = (D.torch().cuda.current_device(), values)
    if key not in _CHOICE_VALUES:
        _CHOICE_VALUES[key] = D.to_dev(np.asarray(values, np.int32))
    return _CHOICE_VALUES[key]


class KDEBank:
    """Per-(set, choice) KDEs for simulations resident on the device.

    rts / choices: device tensors (float32, int8 or int32) in set-major layout
    with int64 ``offsets`` (device tensor, n_sets + 1 entries).
    """

    def __init__(self, rts, choices, offsets, choice_values, *, offsets_host=None, stream=None):
        """offsets_host: the same offsets as a host array, when the caller holds them (saves a
        device-to-host read of them)."""
        t = D.require_gpu()
        self.n_sets = int(offsets.shape[0]) - 1
        self.choice_values = [int(c) for c in choice_values]
        self.C = len(self.choice_values)
        if not 1 <= self.C <= L.MAX_PARTICLES:
            raise ValueError("between 1 and 16 possible choices are supported")
        self.offsets = offsets
        offs_h = offsets.cpu().numpy() if offsets_host is None else np.asarray(offsets_host, np.int64)
        if offs_h.shape != (self.n_sets + 1,):
            raise ValueError("offsets_host must have n_sets + 1 entries")
        self.max_samples = int(np.max(np.diff(offs_h))) if offs_h.size > 1 else 0
        self.d_choice_values = _choice_values_dev(tuple(self.choice_values))
        total = int(rts.shape[0])
        G = self.n_sets * self.C
        self.samples = D.empty((max(total, 1),), t.float32)
        self.kde_count = D.empty((max(G, 1),), t.int32)
        self.choice_total = D.empty((max(G, 1),), t.int32)
        self.bandwidth = D.empty((max(G, 1),), t.float32)
        self.kde_start = D.empty((max(G, 1),), t.int64)
        cb = 1 if choices.dtype == t.int8 else 4
        if choices.dtype not in (t.int8, t.int32):
            raise ValueError("choices must be int8 or int32")
        rc = L.load().ssm_kde_fit(self.n_sets, self.C, D.ptr(self.d_choice_values), D.ptr(rts), D.ptr(choices), cb,
                                  D.ptr(offsets), D.ptr(self.samples), D.ptr(self.kde_count), D.ptr(self.choice_total),
                                  D.ptr(self.bandwidth), D.ptr(self.kde_start), D.stream_ptr(stream))
        L.check(rc, "ssm_kde_fit")

    # ---- sampling -----------------------------------------------------------
    def sample(self, n_kde: int, n_unif: int = 0, *, max_t: float = 20.0, unif_lo: float = 1e-4, seed: int = 0,
               n_by=None, stream=None):
        """KDE resamples (+ uniform points) per set, grouped by choice; returns a point batch."""
        t = D.require_gpu()
        m = int(n_kde) + int(n_unif)
        G = self.n_sets * self.C
        pts = PointBatch(
            rt=D.empty((max(self.n_sets * m, 1),), t.float32),
            choice=D.empty((max(self.n_sets * m, 1),), t.int8),
            grp_begin=D.empty((max(G, 1),), t.int64),
            grp_count=D.empty((max(G, 1),), t.int32),
            m=m,
        )
        d_nby = None if n_by is None else D.to_dev(np.asarray(n_by, np.int32).reshape(-1))
        rc = L.load().ssm_kde_sample(self.n_sets, self.C, int(n_kde), int(n_unif), float(max_t), float(unif_lo),
                                     int(seed) & 0xFFFFFFFFFFFFFFFF, D.ptr(self.samples), D.ptr(self.kde_count),
                                     D.ptr(self.choice_total), D.ptr(self.bandwidth), D.ptr(self.kde_start),
                                     D.ptr(self.offsets), D.ptr(d_nby), D.ptr(pts.rt), D.ptr(pts.choice),
                                     D.ptr(pts.grp_begin), D.ptr(pts.grp_count), D.stream_ptr(stream))
        L.check(rc, "ssm_kde_sample")
        return pts

    # ---- evaluation ---------------------------------------------------------
    def evaluate(self, pts, *, max_t: float = 20.0, lb: float = LB_DEFAULT, eps: float = 10e-5, stream=None):
        """Log-likelihood of every point of `pts` (same indexing as pts.rt)."""
        t = D.require_gpu()
        out = D.empty(tuple(pts.rt.shape), t.float32)
        nch = max(1, -(-self.max_samples // L.KDE_CHUNK))
        ws = D.empty((int(pts.rt.shape[0]) * nch,), t.float32)
        rc = L.load().ssm_kde_eval(self.n_sets, self.C, D.ptr(pts.grp_begin), D.ptr(pts.grp_count), D.ptr(pts.rt),
                                   D.ptr(self.samples), D.ptr(self.kde_count), D.ptr(self.choice_total),
                                   D.ptr(self.bandwidth), D.ptr(self.kde_start), D.ptr(self.offsets),
                                   float(max_t), float(lb), float(eps), self.max_samples, D.ptr(ws), D.ptr(out),
                                   D.stream_ptr(stream))
        L.check(rc, "ssm_kde_eval")
        return out

    def host_state(self):
        """(kde_count, choice_total, bandwidth) as NumPy [n_sets, C] arrays."""
        sh = (self.n_sets, self.C)
        return (self.kde_count.cpu().numpy().reshape(sh), self.choice_total.cpu().numpy().reshape(sh),
                self.bandwidth.cpu().numpy().reshape(sh))


class PointBatch:
    """Evaluation points on the device, grouped by (set, choice)."""

    def __init__(self, rt, choice, grp_begin, grp_count, m):
        self.rt, self.choice, self.grp_begin, self.grp_count, self.m = rt, choice, grp_begin, grp_count, m

    @classmethod
    def from_host(cls, rts: np.ndarray, choice_idx: np.ndarray, n_sets: int, C: int, set_idx=None):
        """Points from host arrays, grouped on the host by (set, choice); returns (batch, order).

        `order[j]` is the input position of the j-th device point, so
        ``out[order] = ll`` restores input order. With n_sets == 1, set_idx may be omitted.
        """
        rts = np.asarray(rts, np.float32).reshape(-1)
        ci = np.asarray(choice_idx, np.int64).reshape(-1)
        si = np.zeros_like(ci) if set_idx is None else np.asarray(set_idx, np.int64).reshape(-1)
        if set_idx is None and n_sets != 1:
            raise ValueError("set_idx is required when n_sets > 1")
        if ci.size and (ci.min() < 0 or ci.max() >= C or si.min() < 0 or si.max() >= n_sets):
            raise ValueError("choice or set index out of range")
        g = si * C + ci
        order = np.argsort(g, kind="stable")
        counts = np.bincount(g, minlength=n_sets * C).astype(np.int32)
        begins = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        D.require_gpu()
        return cls(D.to_dev(rts[order]), D.to_dev(ci[order].astype(np.int8)), D.to_dev(begins), D.to_dev(counts),
                   rts.shape[0]), order


class LogKDE:
    """Drop-in for kde_class.LogKDE: per-choice Gaussian KDE on log RT (GPU)."""

    def __init__(self, simulator_data: dict, bandwidth_type: str = "silverman", auto_bandwidth: bool = True,
                 displace_t: bool = False):
        if bandwidth_type != "silverman":
            raise ValueError(f"Bandwidth type {bandwidth_type} not supported yet")
        self.simulator_info = simulator_data["metadata"]
        self.displace_t = displace_t
        self.auto_bandwidth = auto_bandwidth
        self.bandwidth_type = bandwidth_type
        if displace_t:
            t_vals = np.unique(simulator_data["metadata"]["t"])
            if t_vals.shape[0] != 1:
                raise ValueError("Multiple t values in simulator data. Can't shift.")
            self.displace_t_val = float(t_vals[0])
        data = deepcopy(simulator_data)
        if "rts" not in data and "log_rts" in data:
            lr = np.asarray(data["log_rts"], np.float64)
            data["rts"] = np.where(lr != OMISSION_SENTINEL, np.exp(lr), OMISSION_SENTINEL)
        elif "rts" not in data:
            raise ValueError("simulator_data dictionary must contain either rts or log_rts or both as keys!")
        rts = np.asarray(data["rts"], np.float32).reshape(-1)
        choices = np.asarray(data["choices"]).reshape(-1)
        if displace_t:
            keep = rts != OMISSION_SENTINEL
            rts = np.where(keep, rts - np.float32(self.displace_t_val), rts).astype(np.float32)
        self.choices_list = list(np.unique(self.simulator_info["possible_choices"]))
        self.n_trials = int(choices.shape[0])
        self._bank = KDEBank(D.to_dev(rts), D.to_dev(choices.astype(np.int32)),
                             D.to_dev(np.array([0, rts.shape[0]], np.int64)), self.choices_list)
        cnt, tot, h = self._bank.host_state()
        self.data = {
            "choices": self.choices_list,
            "choice_proportions": [float(x) / max(self.n_trials, 1) for x in tot[0]],
            "n_trials": self.n_trials,
            "n_kde": [int(x) for x in cnt[0]],
        }
        self.bandwidths = [float(b) if b > 0 else "no_base_data" for b in h[0]]

    def compute_bandwidths(self, bandwidth_type: str = "silverman") -> list:
        """Silverman bandwidth per choice, "no_base_data" where a choice has no draws (kde_class.py:89-126)."""
        if bandwidth_type != "silverman":
            raise ValueError(f"Bandwidth type {bandwidth_type} not supported yet")
        return list(self.bandwidths)

    def kde_eval(self, data: dict, log_eval: bool = True, lb: float = LB_DEFAULT, eps: float = 10e-5,
                 filter_rts: float = OMISSION_SENTINEL) -> np.ndarray:
        """Log-likelihood at (rt, choice) pairs (kde_class.py:188-249)."""
        if "rts" in data:
            rts = np.asarray(data["rts"], np.float64).reshape(-1)
        elif "log_rts" in data:
            lr = np.asarray(data["log_rts"], np.float64).reshape(-1)
            rts = np.exp(lr)
            rts[lr == filter_rts] = filter_rts
        else:
            raise ValueError("data dictionary must contain either rts or log_rts as keys!")
        choices = np.asarray(data["choices"]).reshape(-1)
        mask = (np.asarray(data["rts"]).reshape(-1) != filter_rts) if "rts" in data else \
            (np.asarray(data["log_rts"]).reshape(-1) != filter_rts)
        rts, choices = rts[mask], choices[mask]
        if rts.shape != choices.shape:
            raise ValueError("rts and choices need to have matching shapes in data dictionary!")
        if self.displace_t:
            rts = rts - self.displace_t_val
        lookup = {c: i for i, c in enumerate(self.choices_list)}
        try:
            ci = np.array([lookup[c] for c in choices.tolist()], np.int64)
        except KeyError as e:  # same failure as list.index in the reference
            raise ValueError(f"{e.args[0]} is not in list") from None
        if rts.size == 0:
            return np.squeeze(np.zeros((0, 1)))
        pts, order = PointBatch.from_host(rts, ci, 1, len(self.choices_list))
        ll = self._bank.evaluate(pts, max_t=float(self.simulator_info["max_t"]), lb=lb, eps=eps).cpu().numpy()
        out = np.empty(rts.shape[0], np.float64)
        out[order] = ll
        out = out[:, None]
        return np.squeeze(out) if log_eval else np.squeeze(np.exp(out))

    def kde_sample(self, n_samples: int = 2000, use_empirical_choice_p: bool = True,
                   alternate_choice_p: np.ndarray | float = 0.0, random_state: int | None = None) -> dict:
        """Resample (rt, choice) from the KDE (kde_class.py:331-434)."""
        if isinstance(alternate_choice_p, float):
            alt = [alternate_choice_p]
        elif isinstance(alternate_choice_p, np.ndarray):
            alt = alternate_choice_p.tolist()
        else:
            alt = list(alternate_choice_p)
        C = len(self.choices_list)
        if not (len(alt) == C or use_empirical_choice_p):
            raise ValueError("alternate_choice_p must be of the same length as the number of choices")
        n_by = None
        if not use_empirical_choice_p:
            n_by = [round(n_samples * alt[i]) for i in range(C)]
            if sum(n_by) > n_samples:
                n_by[int(np.argmax(n_by))] -= 1
            elif sum(n_by) < n_samples:
                n_by[int(np.argmax(n_by))] += 1
        seed = int(np.random.default_rng(random_state).integers(0, 2**63))
        if n_samples > 1024:
            parts = []
            done = 0
            while done < n_samples:
                k = min(1024, n_samples - done)
                sub = None if n_by is None else _split_counts(n_by, done, k)
                parts.append(self._sample_block(k, seed + done, sub))
                done += k
            rts = np.concatenate([p[0] for p in parts])
            ch = np.concatenate([p[1] for p in parts])
        else:
            rts, ch = self._sample_block(n_samples, seed, n_by)
        if self.displace_t:
            rts = rts + self.displace_t_val
        rts = rts.reshape(-1, 1)
        return {"rts": rts, "log_rts": np.log(rts), "choices": ch.reshape(-1, 1), "metadata": self.simulator_info}

    def _sample_block(self, n, seed, n_by):
        pts = self._bank.sample(n, 0, max_t=float(self.simulator_info["max_t"]), seed=seed, n_by=n_by)
        rts = pts.rt[:n].cpu().numpy().astype(np.float64)
        ci = pts.choice[:n].cpu().numpy().astype(np.int64)
        ch = np.asarray(self.choices_list, np.float64)[ci]
        return rts, ch


def _split_counts(n_by, start, k):
    """Per-choice counts of the points start..start+k of a choice-grouped sequence."""
    edges = np.concatenate([[0], np.cumsum(n_by)])
    out = []
    for i in range(len(n_by)):
        lo, hi = max(edges[i], start), min(edges[i + 1], start + k)
        out.append(max(0, int(hi - lo)))
    return out
