"""Measured compute roofs for the VALU-bound kernels (EM walk, KDE pairs, surrogate hidden units).

The EM and KDE kernels do no MFMA work and move a few bytes per thousand
instructions, so their roof is VALU issue. The probes in probe_kernel.hip run
each kernel's inner loop branch-free at full occupancy; the rate they reach is
the ceiling a kernel with perfect lane utilisation could reach on this chip.
"""

from __future__ import annotations

import numpy as np

from . import _device as D
from . import _lib as L

KINDS = {"em_block": 0, "kde_pair": 1, "mad_u64_u32": 2, "xor_b32": 3, "fma_f32": 4, "log_f32": 5, "exp_f32": 6,
         "sin_f32": 7, "sqrt_f32": 8, "hidden_unit": 9, "lca4_step": 10, "race4_weibull_step": 11}
# work units per lane-iteration: EM steps, KDE pairs, or instructions
UNITS = {0: 4, 1: 16, 2: 8, 3: 8, 4: 8, 5: 8, 6: 8, 7: 8, 8: 8, 9: 8, 10: 1, 11: 1}

# MI355X's published FP32 vector peak, 157.3 TFLOP/s, is one packed FMA (2 lanes x 2 FLOP)
# per lane per quad-cycle on 1024 SIMDs at 2.4 GHz: 39.3e12 unpacked FMA lane-instructions/s.
# Integer and bitwise VALU streams issue faster than one wave-instruction per SIMD per
# quad-cycle on gfx950 (up to 1.12 measured), transcendentals take two quad-cycles.
SIMD_QUAD_CYCLES_PER_S = 256 * 4 * 2.4e9 / 4
PEAK_VALU_LANE_OPS = SIMD_QUAD_CYCLES_PER_S * 64


def probe_rate(kind: str | int, iters: int = 4096, waves_per_simd: int = 8, reps: int = 3) -> float:
    """Units per second (EM steps, KDE pairs or instructions) of a probe at full occupancy."""
    t = D.require_gpu()
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    cus = L.load().ssm_device_cu_count()
    n_threads = cus * 4 * waves_per_simd * 64
    out = D.empty((n_threads,), t.float32)
    lib = L.load()
    L.check(lib.ssm_valu_probe(k, 16, n_threads, D.ptr(out), D.stream_ptr()), "probe")  # warm-up
    best = 0.0
    for _ in range(reps):
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        L.check(lib.ssm_valu_probe(k, iters, n_threads, D.ptr(out), D.stream_ptr()), "probe")
        e.record()
        e.synchronize()
        sec = s.elapsed_time(e) / 1e3
        best = max(best, n_threads * iters * UNITS[k] / sec)
    return best


def instruction_table(iters: int = 2048) -> dict:
    """Measured lane-instructions/s per probed instruction stream and its cost in quad-cycle
    issue slots (one unpacked FP32 FMA per lane per quad-cycle = 1 slot)."""
    out = {}
    for name in ("mad_u64_u32", "xor_b32", "fma_f32", "log_f32", "exp_f32", "sin_f32", "sqrt_f32"):
        r = probe_rate(name, iters)
        out[name] = {"lane_ops_per_s": r, "slots": PEAK_VALU_LANE_OPS / r}
    return out


# ---- workloads for hardware-counter runs ---------------------------------------------------
# `python -m ssmhip.roofline <workload>` runs one kernel family on synthetic inputs of the bench's
# shapes and nothing else, so that a profiler's per-dispatch counters belong to that kernel.

def run_probes(iters: int = 2048, waves_per_simd: int = 8) -> dict:
    """Every probe once at full occupancy, in KINDS order; returns units/s per probe."""
    t = D.require_gpu()
    cus = L.load().ssm_device_cu_count()
    n_threads = cus * 4 * waves_per_simd * 64
    out = D.empty((n_threads,), t.float32)
    lib = L.load()
    rates = {}
    for name, k in KINDS.items():
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        L.check(lib.ssm_valu_probe(k, iters, n_threads, D.ptr(out), D.stream_ptr()), "probe")
        e.record()
        e.synchronize()
        rates[name] = n_threads * iters * UNITS[k] / (s.elapsed_time(e) / 1e3)
    return rates


def run_kde_eval(n_sets: int = 1024, n_samples: int = 100_000, passes: int = 2, seed: int = 1) -> dict:
    """Angle-DDM sets simulated, KDEs fitted, 800 resampled + 200 uniform points per set, then
    `passes` evaluations (kde_eval_partial_kernel + kde_eval_finish_kernel) of those points."""
    from .batched import BatchedSimulation
    from .kde import KDEBank
    from .lan import sample_angle_uniform

    t = D.require_gpu()
    theta = sample_angle_uniform(n_sets, seed=seed)
    sim = BatchedSimulation("angle", theta, n_samples, seed=seed, smooth_unif=True, launch=None)
    sim.run()
    bank = KDEBank(sim.rts, sim.choices, sim.job.offsets, [-1, 1])
    pts = bank.sample(800, 200, seed=seed)
    kc = bank.kde_count.cpu().numpy().reshape(n_sets, 2)
    gc = pts.grp_count.cpu().numpy().reshape(n_sets, 2)
    pairs = float(np.sum(kc.astype(np.float64) * gc))
    ms = []
    for _ in range(passes):
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        bank.evaluate(pts)
        e.record()
        e.synchronize()
        ms.append(s.elapsed_time(e))
    return {"pairs": pairs, "ms": ms, "pairs_per_s": pairs / (min(ms) / 1e3)}


def run_mlp_forward(n_rows: int = 4_000_000, reps: int = 3, seed: int = 0, warmup: int = 0) -> dict:
    """The surrogate's forward (mlp_forward_kernel) on random rows with random weights: `warmup`
    untimed launches, so that the engine clock has settled, then `reps` timed ones."""
    from .surrogate import MLPSurrogate

    t = D.require_gpu()
    net = MLPSurrogate(seed=seed)
    rng = np.random.default_rng(seed)
    lo = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1, 0.0, -1.0])
    hi = np.array([3.0, 3.0, 0.9, 2.0, 1.3, 20.0, 1.0])
    x = D.to_dev(rng.uniform(lo, hi, (n_rows, 7)).astype(np.float32))
    net.set_input_normalisation(x)
    for _ in range(warmup):
        net(x)
    # back-to-back launches between one pair of events: a launch timed alone would also
    # count the host's launch latency while the device waits
    s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        net(x)
    e.record()
    e.synchronize()
    ms = s.elapsed_time(e) / max(reps, 1)
    return {"rows": n_rows, "launches": reps, "ms_per_launch": ms, "rows_per_s": n_rows / (ms / 1e3)}


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python -m ssmhip.roofline", description=__doc__.splitlines()[0])
    ap.add_argument("workload", choices=("probes", "kde-eval", "mlp-forward"))
    ap.add_argument("--iters", type=int, default=2048, help="probe loop iterations")
    ap.add_argument("--sets", type=int, default=1024)
    ap.add_argument("--samples", type=int, default=100_000)
    ap.add_argument("--rows", type=int, default=4_000_000)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=0, help="untimed launches before the timed ones (mlp-forward)")
    a = ap.parse_args(argv)
    if a.workload == "probes":
        res = run_probes(a.iters)
    elif a.workload == "kde-eval":
        res = run_kde_eval(a.sets, a.samples, a.reps)
    else:
        res = run_mlp_forward(a.rows, a.reps, warmup=a.warmup)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
