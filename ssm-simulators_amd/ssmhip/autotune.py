"""Launch-shape autotuner for the persistent EM kernel.

Chooses (block_size, blocks_per_cu, refill_min, queue_block) — threads per
block, resident blocks per CU in the persistent grid, how many idle lanes a
wavefront tolerates before it refills them, and the largest index range a
wavefront claims from the global queue per atomic — by timing candidate
shapes on a job the caller supplies (BatchedSimulation passes a bounded
sample of its own sets, so the sample keeps the workload's mix of short and
long walks). Results are cached per (device, model, boundary, particles, replay mode,
grid length, power-of-two classes of the trajectory and set counts) in memory and in a JSON
file (SSMHIP_TUNE_CACHE, default ~/.cache/ssmhip/tune.json).
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _device as D
from .em import LaunchParams

DEFAULT_CANDIDATES = [
    (256, 8, 8, 64), (256, 8, 4, 64), (256, 8, 2, 64), (256, 8, 4, 32), (256, 8, 8, 32), (256, 8, 8, 128),
    (256, 8, 16, 64), (512, 4, 8, 64), (512, 4, 4, 64), (128, 16, 8, 64), (1024, 2, 8, 64), (256, 8, 4, 128),
    (256, 8, 32, 64), (512, 4, 16, 64), (1024, 2, 16, 64), (512, 4, 32, 64),
    # resident waves at the kernel's 7-per-SIMD occupancy, and longer queue claims
    (256, 7, 4, 128), (512, 3, 4, 128), (512, 3, 4, 256), (256, 8, 4, 256), (512, 4, 4, 128),
]

_memo: dict[str, LaunchParams] = {}


def _cache_path() -> Path:
    return Path(os.environ.get("SSMHIP_TUNE_CACHE", str(Path.home() / ".cache" / "ssmhip" / "tune.json")))


def _key(job) -> str:
    """Cache key: device, model, boundary, particles, replay mode, the grid length (n_steps) and
    power-of-two classes of the trajectory and set counts, so that a shape tuned on a small job
    is not reused for a batch a hundred times larger."""
    t = D.torch()
    name = t.cuda.get_device_name() if t.cuda.is_available() else "cpu"
    p = job.prob
    n_cls = max(0, int(p.n_total)).bit_length()
    k_cls = max(0, int(p.n_sets)).bit_length()
    return (f"{name}|m{p.model}|b{p.boundary}|p{p.n_particles}|r{int(p.replay_stride > 0)}|q"
            f"|s{p.n_steps}|n{n_cls}|k{k_cls}")


def _load_file() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except Exception:
        return {}


def _store(key: str, lp: LaunchParams) -> None:
    data = _load_file()
    data[key] = [lp.block_size, lp.blocks_per_cu, lp.refill_min, lp.queue_block]
    try:
        path = _cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    except OSError:
        pass


def time_launch(job, lp: LaunchParams, reps: int = 3) -> float:
    """Best-of-`reps` time (ms) of a launch, measured with events on the current stream."""
    t = D.torch()
    times = []
    for _ in range(reps):
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        job.launch(lp)
        e.record()
        e.synchronize()
        times.append(s.elapsed_time(e))
    return float(np.min(times))


def tune(job, candidates=None, use_cache: bool = True, verbose: bool = False, reps: int = 3) -> LaunchParams:
    """Pick the fastest launch shape for `job` (an em.EmJob, already on the device)."""
    key = _key(job)
    if use_cache:
        if key in _memo:
            return _memo[key]
        cached = _load_file().get(key)
        if cached:
            lp = LaunchParams(*cached)
            _memo[key] = lp
            return lp
    cands = [LaunchParams(*c) for c in (candidates or DEFAULT_CANDIDATES)]
    job.launch(cands[0])  # warm-up (module load, first-touch)
    D.torch().cuda.synchronize()
    best, best_ms = cands[0], float("inf")
    for lp in cands:
        ms = time_launch(job, lp, reps)
        if verbose:
            print(f"[autotune] block={lp.block_size} per_cu={lp.blocks_per_cu} refill={lp.refill_min}: {ms:.3f} ms",
                  flush=True)
        if ms < best_ms:
            best, best_ms = lp, ms
    _memo[key] = best
    if use_cache:
        _store(key, best)
    return best
