"""Auxiliary labels of the LAN data generator (test oracle).

Follows SimulationPipeline._compute_auxiliary_labels
(ssms/dataset_generators/pipelines/simulation_pipeline.py:258-351) for one
parameter set, and the float32 concatenation of TrainingDataGenerator
(ssms/dataset_generators/lan_mlp.py:533-541) over sets.
The histograms use np.histogram with explicit np.linspace edges, the same NumPy
calls the reference makes.
Also the simulation filter's per-choice statistics and keep rule
(simulation_pipeline.py:199-256).
Pinned by: a hand-built case whose counts are written out by hand in
tests/test_oracle_aux.py (rts exactly on bin edges, at max_t, above max_t,
omitted, choices outside possible_choices); the filter statistics by a
hand-written case (ties for the mode, an empty choice, fewer than 5 values).
"""

from __future__ import annotations

import numpy as np

OMISSION = -999.0


def labels_one_set(rts, choices, possible_choices, max_t=20.0) -> dict:
    """The reference's auxiliary label dict for one set (float64, n_trials = 1)."""
    rts = np.asarray(rts, np.float32).reshape(-1)
    choices = np.asarray(choices).reshape(-1)
    C = len(possible_choices)
    n = len(choices)
    choice_p = np.zeros((1, C))
    for i, c in enumerate(possible_choices):  # :285-288
        choice_p[0, i] = np.sum(choices == c) / n
    non_omitted = choices != OMISSION  # :291-301
    choice_p_no_omission = np.zeros((1, C))
    if np.any(non_omitted):
        cno = choices[non_omitted]
        for i, c in enumerate(possible_choices):
            choice_p_no_omission[0, i] = np.sum(cno == c) / len(cno)
    else:
        choice_p_no_omission[0, :] = 1.0 / C
    omission_p = np.array([[np.sum(choices == OMISSION) / n]])  # :304
    mx = max(possible_choices)  # :306-312
    nogo_p = np.array([[np.sum((choices != mx) | (rts == OMISSION)) / n]])
    go_p = 1 - nogo_p
    b128 = np.linspace(0, max_t, 129)  # :314-331
    b256 = np.linspace(0, max_t, 257)
    h128 = np.zeros((1, 128, C))
    h256 = np.zeros((1, 256, C))
    for i, c in enumerate(possible_choices):
        r = rts[choices == c]
        r = r[r != OMISSION]
        if len(r) > 0:
            h128[0, :, i], _ = np.histogram(r, bins=b128)
            h256[0, :, i], _ = np.histogram(r, bins=b256)
    return {"choice_p": choice_p, "choice_p_no_omission": choice_p_no_omission, "omission_p": omission_p,
            "nogo_p": nogo_p, "go_p": go_p, "binned_128": h128, "binned_256": h256}


def labels_all_sets(rts, choices, offsets, possible_choices, max_t=20.0) -> dict:
    """Per-set labels stacked and cast as TrainingDataGenerator does: cpn_labels is the
    probability of the second choice for two-choice models (simulation_pipeline.py:142-151)."""
    per = [labels_one_set(rts[offsets[k]:offsets[k + 1]], choices[offsets[k]:offsets[k + 1]], possible_choices,
                          max_t) for k in range(len(offsets) - 1)]
    two = len(possible_choices) == 2

    def cat(key, fn=lambda a: a):
        return np.concatenate([fn(d[key]) for d in per]).astype(np.float32)

    return {
        "cpn_labels": cat("choice_p", (lambda a: a[0, 1:2].reshape(1, 1)) if two else (lambda a: a)),
        "cpn_no_omission_labels": cat("choice_p_no_omission",
                                      (lambda a: a[0, 1:2].reshape(1, 1)) if two else (lambda a: a)),
        "opn_labels": cat("omission_p"),
        "gonogo_labels": cat("nogo_p"),
        "binned_128": cat("binned_128"),
        "binned_256": cat("binned_256"),
    }


def filter_stats_one_set(rts, choices, possible_choices) -> np.ndarray:
    """Per-choice (mode, mean, std, mode_cnt_rel, count) of SimulationPipeline._is_valid_simulation
    (simulation_pipeline.py:218-238): over the rts of each choice without omissions,
    scipy.stats.mode (smallest of the most frequent values), np.mean / np.std of the float32
    values, mode count / n (0 below 5 values); (-1, -1, 1, 0, 0) for a choice without draws."""
    from scipy.stats import mode

    rts = np.asarray(rts, np.float32).reshape(-1)
    choices = np.asarray(choices).reshape(-1)
    out = np.zeros((len(possible_choices), 5))
    for i, c in enumerate(possible_choices):
        x = rts[(choices == c) & (rts != np.float32(OMISSION))]
        n = len(x)
        if n == 0:
            out[i] = (-1, -1, 1, 0, 0)
            continue
        m, cnt = mode(x, keepdims=False)
        out[i] = (m, np.mean(x), np.std(x), 0 if n < 5 else cnt / n, n)
    return out


def filter_keep(stats: np.ndarray, filters: dict) -> bool:
    """The keep rule over the choices' statistics (simulation_pipeline.py:239-246); thresholds
    default to keep-everything when a filter key is absent."""
    keep = True
    for mode_, mean_, std_, rel_, n_ in stats:
        keep = keep & (mode_ <= filters.get("mode", float("inf"))) & (mean_ <= filters.get("mean_rt", float("inf"))) \
            & (std_ >= filters.get("std", 0.0)) & (rel_ <= filters.get("mode_cnt_rel", 1.0)) \
            & (n_ >= filters.get("choice_cnt", 0))
    return bool(keep)
