"""Measured compute roofs for the VALU-bound kernels (EM walk, KDE pairs, surrogate hidden units).

The EM and KDE kernels do no MFMA work and move a few bytes per thousand
instructions, so their roof is VALU issue. The probes in probe_kernel.hip run
each kernel's inner loop branch-free at full occupancy; the rate they reach is
the ceiling a kernel with perfect lane utilisation could reach on this chip.
"""

from __future__ import annotations

from . import _device as D
from . import _lib as L

KINDS = {"em_block": 0, "kde_pair": 1, "mad_u64_u32": 2, "xor_b32": 3, "fma_f32": 4, "log_f32": 5, "exp_f32": 6,
         "sin_f32": 7, "sqrt_f32": 8, "hidden_unit": 9, "lca4_step": 10, "race4_weibull_step": 11}
# work units per lane-iteration: EM steps, KDE pairs, or instructions
UNITS = {0: 4, 1: 16, 2: 8, 3: 8, 4: 8, 5: 8, 6: 8, 7: 8, 8: 8, 9: 8, 10: 1, 11: 1}

# MI355X's published FP32 vector peak, 157.3 TFLOP/s, as FMA lane-operations/s (an FMA is
# 2 FLOP); the EM kernel was measured issuing 48e12 lane-level VALU instructions/s, so a
# 16-lane-per-SIMD reading of the chip (39.3e12/s) would understate it
PEAK_VALU_LANE_OPS = 157.3e12 / 2


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
    """Measured lane-instructions/s per probed instruction and its cost in slots of the chip's
    published FP32 FMA rate."""
    out = {}
    for name in ("mad_u64_u32", "xor_b32", "fma_f32", "log_f32", "exp_f32", "sin_f32", "sqrt_f32"):
        r = probe_rate(name, iters)
        out[name] = {"lane_ops_per_s": r, "slots": PEAK_VALU_LANE_OPS / r}
    return out
