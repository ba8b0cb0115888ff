"""Float64 restatement of LogKDE (test oracle).

Follows ssms/support_utils/kde_class.py:
  fit   _attach_data_from_simulator :438-504 (split by possible choice, drop
        -999, p_c = count_c / N incl. omissions), bandwidth_silverman :508-553,
        fit on log(rt) of the float32 rts (:174-176)
  sample counts  kde_sample :387-399 (n_c = round(n p_c) on the float64 product, +-1 on the
        first largest count when the counts do not add up to n)
  eval  __kde_eval_ :251-329 with KernelDensity.score_samples
        (sklearn/neighbors/_kde.py:248-284): exact Gaussian log-sum, no tree
        approximation (rtol = atol = 0 in the reference makes it exact).
The log-sum uses a max shift in float64 (NumPy, or the OpenMP routine of
cpu_baseline.cpp for large inputs).
Pinned by: scikit-learn's KernelDensity (an independent implementation of the
same estimator) on small inputs, the reference tests' bandwidth value
(tests/test_kde_class.py:31-36: constant sample -> ~0.0008) and closed forms
(one sample -> a Gaussian density; floors).
"""

from __future__ import annotations

import numpy as np

OMISSION = -999.0
LOG_2PI_HALF = 0.5 * np.log(2 * np.pi)


def bandwidth_silverman(sample, std_cutoff=1e-3, std_n_1=10.0):
    s = np.asarray(sample, np.float64).reshape(-1)
    n = s.size
    if n > 1:
        std = np.std(s)
        if std < std_cutoff:
            std = std_cutoff
    else:
        std = std_n_1
    return np.power(4 / (3 * n), 1 / 5) * std


def logsumexp_gauss(x: np.ndarray, samples: np.ndarray, h: float) -> np.ndarray:
    """log sum_i exp(-(x - s_i)^2 / 2h^2) for each x (float64, max-shifted)."""
    x = np.asarray(x, np.float64).reshape(-1)
    s = np.asarray(samples, np.float64).reshape(-1)
    if x.size * s.size > 2e7:
        from . import cpu_baseline

        return cpu_baseline.kde_logsum(x, s, h)
    e = -((x[:, None] - s[None, :]) ** 2) / (2 * h * h)
    m = e.max(axis=1)
    return m + np.log(np.exp(e - m[:, None]).sum(axis=1))


def n_by_choice(n_samples: int, p) -> list[int]:
    """round(n * p_c) per choice, Python's round of the float64 product (half to even on that
    value); then, if the counts do not add up to n, the first largest count moves by one
    (kde_class.py:387-399)."""
    n_by = [round(n_samples * float(pc)) for pc in p]
    if sum(n_by) > n_samples:
        n_by[int(np.argmax(n_by))] -= 1
    elif sum(n_by) < n_samples:
        n_by[int(np.argmax(n_by))] += 1
    return n_by


class KDEFit:
    def __init__(self, rts, choices, possible_choices):
        rts = np.asarray(rts, np.float32).reshape(-1)
        choices = np.asarray(choices).reshape(-1)
        self.N = rts.shape[0]
        self.choices = list(np.unique(possible_choices))
        self.logs, self.p, self.h = [], [], []
        for c in self.choices:
            r = rts[choices == c]
            self.p.append(r.shape[0] / self.N)
            keep = r[r != OMISSION]
            lr = np.log(keep).astype(np.float64)  # np.log of float32 data, as the reference
            self.logs.append(lr)
            self.h.append(bandwidth_silverman(lr) if lr.size else None)

    def n_by_choice(self, n_samples: int) -> list[int]:
        """Per-choice KDE sample counts of kde_sample (kde_class.py:387-399)."""
        return n_by_choice(n_samples, self.p)

    def eval(self, rts, choices, max_t=20.0, lb=-66.774, eps=10e-5):
        rts = np.asarray(rts, np.float64).reshape(-1)
        choices = np.asarray(choices).reshape(-1)
        out = np.zeros(rts.shape[0])
        lr = np.log(np.maximum(rts, eps))
        for i, c in enumerate(self.choices):
            m = choices == c
            if not np.any(m):
                continue
            if self.h[i] is None:
                out[m] = -np.log(self.N * max_t)
                continue
            n, h = self.logs[i].size, self.h[i]
            ll = np.log(self.p[i]) + logsumexp_gauss(lr[m], self.logs[i], h) - np.log(n) - np.log(h) - LOG_2PI_HALF - lr[m]
            ll[(ll <= lb) | (rts[m] <= 0)] = lb
            out[m] = ll
        return out
