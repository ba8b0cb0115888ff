"""Benchmark: batched angle-DDM trajectories/s on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): 4096 angle-DDM
parameter sets drawn uniformly from v in [-3,3], a in [0.3,3], z in [0.1,0.9],
t in [0.001,2], theta in [-0.1,1.3]; per-set n_samples = floor(10^U(3,5));
dt = 0.001, max_t = 20, float32 rts, int8 choices; all sets in ONE persistent
kernel launch replayed from a HIP graph, launch shape chosen by the autotuner.

Prints one JSON line (see the repository documentation for the fields).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "ssm-simulators_amd"))
sys.path.insert(0, str(REPO))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_leg(theta, ns, target_seconds: float = 15.0):
    """Time the C++/OpenMP restatement (MT19937 + Ziggurat) on a bounded sample of the same workload."""
    from oracle import cpu_baseline
    from oracle import em_numpy as E

    cpu_baseline.load()
    cores = cpu_baseline.n_cores()
    n_sets = theta.shape[0]
    v, a, z, t, th = (theta[:, i] for i in range(5))
    dt = 0.001

    def run(idx, scale):
        nsub = np.maximum(1, np.floor(ns[idx] * scale)).astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(nsub)]).astype(np.int64)
        table = E.boundary_table("angle", {"a": a[idx], "theta": th[idx]}, idx.size, 20.0, dt)
        dl = np.minimum(np.float32(20.0), (np.float32(999.0) - t[idx]).astype(np.float32))
        t0 = time.perf_counter()
        _, _, steps = cpu_baseline.em_simulate(0, v[idx], z[idx], np.full(idx.size, E.sqrt_dt32(dt), np.float32),
                                               t[idx], np.full(idx.size, 999.0, np.float32), dl, table, offs, dt=dt,
                                               seed=1, n_threads=cores)
        return time.perf_counter() - t0, int(offs[-1]), steps

    idx = np.arange(0, n_sets, 4)  # every 4th set keeps the parameter mix
    sec, n, _ = run(idx, 0.01)
    rate = n / max(sec, 1e-9)
    scale = min(1.0, 0.01 * target_seconds / max(sec, 1e-6))
    sec, n, steps = run(idx, scale)
    return {
        "value": n / sec,
        "unit": "trajectories/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{idx.size} of {n_sets} sets (every 4th), n_samples scaled by {scale:.4f}: {n} trajectories, "
                  f"{steps} EM steps in {sec:.2f} s ({steps / sec:.3e} steps/s)",
        "calibration_rate": rate,
    }


def sample_accumulator_sets(model: str, n_sets: int = 1024, seed: int = 3):
    """Config-3 parameter sets: LCA-4 (v_i in [0,2.5], a in [1,3], z_i in [0,0.5], g,b in [-1,1], t in [0,2])
    or race-4 with a Weibull-collapsing bound (race_4 bounds, alpha in [0.3,5], beta in [0.3,7])."""
    rng = np.random.default_rng(seed)
    if model == "lca_4":
        lo = [0.0] * 4 + [1.0] + [0.0] * 4 + [-1.0, -1.0, 0.0]
        hi = [2.5] * 4 + [3.0] + [0.5] * 4 + [1.0, 1.0, 2.0]
    else:
        lo = [0.0] * 4 + [1.0] + [0.0] * 4 + [0.0, 0.3, 0.3]
        hi = [2.5] * 4 + [3.0] + [0.9] * 4 + [2.0, 5.0, 7.0]
    return rng.uniform(lo, hi, size=(n_sets, len(lo))).astype(np.float32)


def time_graph(sim, steps: int, warmup: int):
    import torch

    sim.capture()
    for _ in range(warmup):
        sim.replay()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        sim.replay()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / steps


def queue_order(args):
    return None if args.queue_order == "index" else args.queue_order


def accumulator_leg(args):
    """LCA-4 and race-4/Weibull, 1024 sets x 20,000 samples each, against the CPU restatement."""
    from oracle import cpu_baseline
    from oracle import em_numpy as E
    from ssmhip import batched, roofline

    out = {}
    for model, kind in (("lca_4", 2), ("race_4_weibull", 1)):
        theta = sample_accumulator_sets(model)
        sim = batched.BatchedSimulation(model, theta, 20000, seed=3, launch="auto", tune_sample=None,
                                        queue_order=queue_order(args))
        ms = time_graph(sim, max(2, args.steps), 1)
        steps = sim.steps()
        lp = sim.launch_params
        rec = {"trajectories": sim.n_total, "ms_per_launch": ms, "trajectories_per_s": sim.n_total / (ms / 1e3),
               "em_steps_per_s": steps / (ms / 1e3),
               "launch": [lp.block_size, lp.blocks_per_cu, lp.refill_min, lp.queue_block],
               "lane_utilisation": steps / max(1, sim.lane_slots())}
        if not args.no_probe:
            # roof: the same model's step (bound test, one Philox block, 4-accumulator update)
            # branch-free at full occupancy
            peak = roofline.probe_rate("lca4_step" if model == "lca_4" else "race4_weibull_step", iters=2048)
            rec["roofline"] = {"bound": "valu", "achieved": rec["em_steps_per_s"] / 1e9, "peak": peak / 1e9,
                               "unit": "G EM-steps/s", "frac": rec["em_steps_per_s"] / peak}
        if not args.no_cpu:
            idx = np.arange(0, theta.shape[0], 32)
            n_cpu = 20000
            v = theta[idx, 0:4]
            z = theta[idx, 5:9]
            a = theta[idx, 4]
            if model == "lca_4":
                g, b, t = theta[idx, 9], theta[idx, 10], theta[idx, 11]
                table = E.boundary_table("constant", {"a": a}, idx.size, 20.0, 0.001)
            else:
                g = b = None
                t = theta[idx, 9]
                table = E.boundary_table("weibull_cdf", {"a": a, "alpha": theta[idx, 10], "beta": theta[idx, 11]},
                                         idx.size, 20.0, 0.001)
            offs = (np.arange(idx.size + 1) * n_cpu).astype(np.int64)
            dl = np.minimum(np.float32(20.0), (np.float32(999.0) - t).astype(np.float32))
            t0 = time.perf_counter()
            cpu_baseline.em_simulate(kind, v, z, np.full((idx.size, 4), E.sqrt_dt32(0.001), np.float32), t,
                                     np.full(idx.size, 999.0, np.float32), dl, table, offs, dt=0.001, seed=1, g=g,
                                     b=b, n_threads=cpu_baseline.n_cores())
            sec = time.perf_counter() - t0
            rec["cpu_baseline"] = {"value": idx.size * n_cpu / sec, "unit": "trajectories/s",
                                   "cores": cpu_baseline.n_cores(), "kind": "port",
                                   "sample": f"{idx.size} of 1024 sets (every 32nd) x {n_cpu} samples in {sec:.2f} s"}
            rec["gpu_over_cpu"] = rec["trajectories_per_s"] / rec["cpu_baseline"]["value"]
        out[model] = rec
        log(f"accumulators {model}: {rec['trajectories_per_s']:.3e} traj/s"
            + (f", {rec['roofline']['frac'] * 100:.0f}% of the step roof" if "roofline" in rec else "")
            + (f", {rec['gpu_over_cpu']:.0f}x CPU" if "gpu_over_cpu" in rec else ""))
    return out


def ddm_leg(args):
    """Plain DDM through the drop-in front end: Simulator('ddm').simulate with v=1, a=1.5, z=0.5,
    t=0.2, 100,000 samples, seed 0 (dt 1e-3, max_t 20, smooth_unif, full metadata), end to end:
    parameter preprocessing, the launch, the copy of rts and choices to the host, the trajectory
    of trial 0 and the boundary of the metadata. Median of 5 calls after one warm-up call."""
    import torch

    from oracle import cpu_baseline
    from oracle import em_numpy as E
    from ssmhip.simulator import Simulator

    theta = {"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.2}
    sim = Simulator("ddm")
    sim.simulate(theta, n_samples=100_000, random_state=0)
    times = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = sim.simulate(theta, n_samples=100_000, random_state=0)
        times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    p_up = float(np.mean(out["choices"] == 1))
    rec = {"workload": "Simulator('ddm').simulate, v=1, a=1.5, z=0.5, t=0.2, 1e5 samples, seed 0",
           "seconds_end_to_end": sec, "trajectories_per_s": 100_000 / sec, "p_upper": p_up,
           "mean_rt": float(np.mean(out["rts"]))}
    if not args.no_cpu:
        table = E.boundary_table("constant", {"a": [1.5]}, 1, 20.0, 0.001)
        t0 = time.perf_counter()
        cpu_baseline.em_simulate(0, [1.0], [0.5], [E.sqrt_dt32(0.001)], [0.2], [999.0], [20.0], table,
                                 np.array([0, 100_000], np.int64), dt=0.001, seed=0, smooth_unif=True,
                                 n_threads=cpu_baseline.n_cores())
        csec = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": 100_000 / csec, "unit": "trajectories/s", "cores": cpu_baseline.n_cores(),
                               "kind": "port", "sample": f"the whole workload, 1e5 trajectories in {csec:.3f} s"}
        rec["gpu_over_cpu"] = rec["trajectories_per_s"] / rec["cpu_baseline"]["value"]
    log(f"ddm (1 set x 1e5, end to end): {sec * 1e3:.2f} ms, {rec['trajectories_per_s']:.3e} traj/s, "
        f"P(upper) {p_up:.4f}" + (f", {rec['gpu_over_cpu']:.0f}x CPU" if "gpu_over_cpu" in rec else ""))
    return rec


def hssm_leg(args):
    """HSSM posterior-predictive draws: 1000 posterior draws x 1000 data trials of the angle model
    (per-trial drift), one sample each, through hssm_support.rng_fn; the device-only rate of the
    same 1e6 one-sample sets from a replayed graph; the CPU restatement of the reference's
    parallel path (per-trial boundary precompute + walk) on a sample of the rows."""
    import torch

    from oracle import cpu_baseline
    from oracle import em_numpy as E
    from ssmhip import batched
    from ssmhip import hssm_support as H

    draws, trials = 1000, 1000
    rng = np.random.default_rng(17)
    lo = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    hi = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    post = rng.uniform(lo, hi, size=(draws, 5)).astype(np.float32)
    x = rng.uniform(-1, 1, trials).astype(np.float32)  # a per-trial covariate of the drift
    v = np.clip(post[:, :1] + 0.5 * x[None, :], -3, 3).astype(np.float32)
    args_ = [v] + [post[:, j:j + 1] for j in range(1, 5)]
    fn = H.get_simulator_fun_internal("angle")
    # steady state: one call of the full size first (module load, device buffers), then the
    # median of three timed calls
    H.rng_fn(args_, None, np.random.default_rng(0), fn, 2)
    torch.cuda.synchronize()
    walls = []
    for rep in range(3):
        t0 = time.perf_counter()
        out = H.rng_fn(args_, None, np.random.default_rng(1 + rep), fn, 2)
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    rows = draws * trials
    assert out.shape == (draws, trials, 2)
    theta = np.column_stack([np.broadcast_to(a, (draws, trials)).reshape(-1) for a in args_]).astype(np.float32)
    sim = batched.BatchedSimulation("angle", theta, 1, seed=2, launch="auto", tune_sample=None,
                                    queue_order=queue_order(args))
    ms = time_graph(sim, max(2, args.steps), 1)
    rec = {"workload": "angle, 1000 posterior draws x 1000 trials, 1 sample each", "rows": rows,
           "rows_per_s_end_to_end": rows / wall, "end_to_end_s": wall,
           "device_rows_per_s": rows / (ms / 1e3), "device_ms": ms}
    if not args.no_cpu:
        # every 5th row, in chunks of 5000 (the per-trial boundary tables of a chunk take 400 MB)
        idx = np.arange(0, rows, 5)
        sec = 0.0
        for c0 in range(0, idx.size, 5000):
            th = theta[idx[c0:c0 + 5000]]
            K = th.shape[0]
            t1 = time.perf_counter()
            table = E.boundary_table("angle", {"a": th[:, 1], "theta": th[:, 4]}, K, 20.0, 0.001)
            dl = np.minimum(np.float32(20.0), (np.float32(999.0) - th[:, 3]).astype(np.float32))
            cpu_baseline.em_simulate(0, th[:, 0], th[:, 2], np.full(K, E.sqrt_dt32(0.001), np.float32), th[:, 3],
                                     np.full(K, 999.0, np.float32), dl, table, np.arange(K + 1, dtype=np.int64),
                                     dt=0.001, seed=1, n_threads=cpu_baseline.n_cores())
            sec += time.perf_counter() - t1
        rec["cpu_baseline"] = {"value": idx.size / sec, "unit": "rows/s", "cores": cpu_baseline.n_cores(),
                               "kind": "port",
                               "sample": f"{idx.size} of the 1e6 rows (every 5th): NumPy boundary per trial + "
                                         f"C++/OpenMP walk, {sec:.2f} s"}
        rec["gpu_over_cpu_end_to_end"] = rec["rows_per_s_end_to_end"] / rec["cpu_baseline"]["value"]
    log(f"hssm rng_fn: {rows / wall:.3e} rows/s end to end ({wall:.2f} s), device {rows / (ms / 1e3):.3e} rows/s"
        + (f", {rec['gpu_over_cpu_end_to_end']:.0f}x CPU" if "gpu_over_cpu_end_to_end" in rec else ""))
    return rec


def lan_leg(args):
    """KDE-labelled LAN rows: 1024 angle sets x 1e5 simulations -> 800 KDE + 200 uniform rows per set."""
    import torch

    from oracle import cpu_baseline, kde_ref
    from oracle import em_numpy as E
    from ssmhip import roofline
    from ssmhip.lan import LanKDELabeller, sample_angle_uniform

    theta = sample_angle_uniform(1024, seed=1)
    # mixture [0.8, 0.2, 0.0] of 1000 rows: 800 KDE resamples + 200 uniform rows, no negative rows
    lab = LanKDELabeller("angle", theta, n_samples=100_000, n_kde=800, n_unif=200, n_neg=0, seed=4, launch="auto",
                         tune_sample=None)
    lab.run(timed=True)  # warm-up
    torch.cuda.synchronize()
    reps = max(2, min(args.steps, 5))
    tot = []
    for _ in range(reps):
        rows = lab.run(timed=True)
        tot.append(lab.times)
    torch.cuda.synchronize()
    # the mean of the timed passes, stage by stage
    from ssmhip.lan import StageTimes

    st = StageTimes(*(float(np.mean([getattr(x, f) for x in tot])) for f in
                      ("simulate_ms", "fit_ms", "sample_ms", "eval_ms", "assemble_ms")))
    n_rows = int(rows.shape[0])
    cnt = lab.bank.kde_count.cpu().numpy().astype(np.int64)
    gcount = lab.points.grp_count.cpu().numpy().astype(np.int64)
    pairs = int(np.sum(cnt * gcount))
    rec = {"rows": n_rows, "ms_per_pass": st.total_ms, "rows_per_s": n_rows / (st.total_ms / 1e3),
           "stage_ms": {"simulate": st.simulate_ms, "kde_fit": st.fit_ms, "kde_sample": st.sample_ms,
                        "kde_eval": st.eval_ms, "assemble": st.assemble_ms},
           "kde_evals_per_s": n_rows / (st.eval_ms / 1e3), "kde_pairs": pairs,
           "kde_pairs_per_s": pairs / (st.eval_ms / 1e3)}
    if not args.no_probe:
        roof = roofline.probe_rate("kde_pair", iters=2048)
        exp_rate = roofline.probe_rate("exp_f32", iters=2048)
        rec["kde_eval_roofline"] = {"bound": "valu (exp)", "achieved": rec["kde_pairs_per_s"] / 1e12,
                                    "peak": roof / 1e12, "unit": "T pairs/s", "frac": rec["kde_pairs_per_s"] / roof,
                                    # one v_exp_f32 per pair: the transcendental rate against the
                                    # probe-measured single-instruction v_exp_f32 rate
                                    "exp_per_s": rec["kde_pairs_per_s"], "exp_probe_per_s": exp_rate,
                                    "exp_frac_of_probe": rec["kde_pairs_per_s"] / exp_rate}
    tpp = counter_value("trans_lane_instructions_per_pair", "kde_eval_lan1024_counters.md") or 1.0
    rec["kde_eval_published_peak"] = {
        "transcendental": {"unit": "T lane-instructions/s", "achieved": tpp * rec["kde_pairs_per_s"] / 1e12,
                           "peak": PEAK_TRANS_PER_S / 1e12, "frac": tpp * rec["kde_pairs_per_s"] / PEAK_TRANS_PER_S,
                           "basis": f"{tpp:.3f} v_exp_f32 per pair (counters); ceiling 2 quad-cycles per "
                                    "transcendental (profiles/valu_issue_probes_counters.md)"},
        "fp32": {"unit": "TFLOP/s", "achieved": 4 * rec["kde_pairs_per_s"] / 1e12, "peak": PEAK_FP32_FLOPS / 1e12,
                 "frac": 4 * rec["kde_pairs_per_s"] / PEAK_FP32_FLOPS, "basis": "sub, fma, add = 4 FLOP per pair"},
        "valu_pipe_busy_counters": 0.960,
    }
    if not args.no_cpu:
        # CPU: per set, simulate 1e5 (MT19937 + Ziggurat, all cores) + exact float64 KDE at the same 1000 points
        pts_rt = lab.points.rt.cpu().numpy().reshape(1024, -1)
        pts_ch = lab.points.choice.cpu().numpy().reshape(1024, -1)
        nsets = 2
        t0 = time.perf_counter()
        for k in range(nsets):
            v, a, z, t, th = theta[k]
            table = E.boundary_table("angle", {"a": [a], "theta": [th]}, 1, 20.0, 0.001)
            rts, ch, _ = cpu_baseline.em_simulate(0, [v], [z], [E.sqrt_dt32(0.001)], [t], [999.0],
                                                  [min(20.0, 999.0 - t)], table, np.array([0, 100_000], np.int64),
                                                  dt=0.001, seed=k, smooth_unif=True,
                                                  n_threads=cpu_baseline.n_cores())
            fit = kde_ref.KDEFit(rts, ch, [-1, 1])
            fit.eval(pts_rt[k].astype(np.float64), np.array([-1, 1])[pts_ch[k]])
        sec = (time.perf_counter() - t0) / nsets
        rec["cpu_baseline"] = {"value": lab.rows_per_set / sec, "unit": "rows/s", "cores": cpu_baseline.n_cores(),
                               "kind": "port",
                               "sample": f"{nsets} sets: 1e5 simulations + 1000 exact float64 KDE evaluations each, "
                                         f"{sec:.2f} s per set"}
        rec["gpu_over_cpu"] = rec["rows_per_s"] / rec["cpu_baseline"]["value"]
    log(f"lan rows: {rec['rows_per_s']:.3e} rows/s, KDE {rec['kde_evals_per_s']:.3e} evals/s")
    return rec, lab, rows


def surrogate_leg(args, lab, kde_evals_per_s, config4_rows=None):
    """Surrogate: label rows of --surrogate-sets angle sets, train, then 4M-row inference against
    the GPU KDE and accuracy against the analytic DDM density."""
    import torch

    from oracle import ddm_analytic
    from ssmhip import roofline
    from ssmhip.surrogate import MLPSurrogate

    from ssmhip.lan import LanKDELabeller, sample_angle_uniform

    # training rows from the same pipeline at a larger simulation budget per set (1e6 draws:
    # the KDE labels' own error falls with the bandwidth), angle sets outside the LAN-rows sample
    parts = []
    done = 0
    torch.cuda.synchronize()
    t_lab = time.perf_counter()
    while done < args.surrogate_sets:
        k = min(2048, args.surrogate_sets - done)
        more = LanKDELabeller("angle", sample_angle_uniform(k, seed=100 + done), n_samples=args.surrogate_samples,
                              n_kde=800, n_unif=200, n_neg=0,
                              seed=200 + done, launch=lab.sim.launch_params)
        parts.append(more.run())
        del more
        done += k
    rows = torch.cat(parts, 0).contiguous()
    del parts
    torch.cuda.synchronize()
    label_s = time.perf_counter() - t_lab
    n = rows.shape[0]
    perm = torch.randperm(n, device=rows.device, generator=torch.Generator(device=rows.device).manual_seed(0))
    val = rows.index_select(0, perm[: n // 50])
    train = rows.index_select(0, perm[n // 50:])
    del rows, perm
    net = MLPSurrogate(seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.fit(train, epochs=args.surrogate_epochs, batch=8192, lr=3e-3, lr_final=5e-5, val_rows=val)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    rng = np.random.default_rng(5)
    N = 4_000_000
    lo = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    hi = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    x = np.concatenate([rng.uniform(lo, hi, (N, 5)), rng.uniform(0, 20, (N, 1)), rng.choice([-1.0, 1.0], (N, 1))],
                       axis=1).astype(np.float32)
    dx = torch.from_numpy(x).cuda()
    for _ in range(150):  # steady state: after idle host-side work the engine clock takes ~100 launches (~35 ms) to settle
        net(dx)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    s.record()
    for _ in range(reps):
        net(dx)
    e.record()
    e.synchronize()
    ms = s.elapsed_time(e) / reps
    evals = N / (ms / 1e3)
    flops = 2 * (7 * 100 + 100 * 100 + 100 * 120 + 120) * N / (ms / 1e3)
    # accuracy at theta = 0 (angle == plain DDM): rows drawn from the model itself
    k = 2000
    th = rng.uniform(lo[:4], hi[:4], (k, 4)).astype(np.float32)
    from ssmhip import batched

    sim = batched.BatchedSimulation("ddm", th, 50, seed=9, launch=None)
    sim.run()
    rts = sim.rts.cpu().numpy().reshape(k, 50)
    chs = sim.choices.cpu().numpy().reshape(k, 50)
    xs, ref = [], []
    for i in range(k):
        m = rts[i] > 0
        v, a, z, t = th[i]
        ll = ddm_analytic.log_density_rt(rts[i][m], chs[i][m], float(v), float(a), float(z), float(t))
        ok = np.isfinite(ll)
        xi = np.zeros((int(ok.sum()), 7), np.float32)
        xi[:, :4] = th[i]
        xi[:, 4] = 0.0
        xi[:, 5] = rts[i][m][ok]
        xi[:, 6] = chs[i][m][ok]
        xs.append(xi)
        ref.append(ll[ok])
    xs_parts = xs
    xs = np.concatenate(xs)
    ref = np.concatenate(ref)
    dxs = torch.from_numpy(xs).cuda()
    pred = net(dxs).cpu().numpy()
    mae = float(np.mean(np.abs(pred - ref)))
    # the same network trained on the LAN leg's own rows (1024 sets x 1e5 simulations, 800 KDE +
    # 200 uniform rows per set: the labelling workload as configured, 1,024,000 rows). Its labels
    # are noisier (1e5 draws per KDE, 32x fewer sets): a Huber delta of 0.1 (near the absolute
    # error) and a larger first learning rate fit them better (0.080 -> 0.063 nats over 3 seeds;
    # the larger label set above is best with its own settings, 0.046 against 0.049)
    mae_c4 = None
    c4_fit = {"epochs": args.surrogate_epochs, "batch": 8192, "lr": 1e-2, "lr_final": 5e-5, "delta": 0.1}
    if config4_rows is not None:
        n4 = config4_rows.shape[0]
        perm4 = torch.randperm(n4, device="cuda", generator=torch.Generator(device="cuda").manual_seed(1))
        net4 = MLPSurrogate(seed=0)
        net4.fit(config4_rows.index_select(0, perm4[n4 // 50:]), val_rows=config4_rows.index_select(0, perm4[: n4 // 50]),
                 **c4_fit)
        mae_c4 = float(np.mean(np.abs(net4(dxs).cpu().numpy() - ref)))
        del net4, perm4
    # the KDE labels' own error at the same points (simulated KDE per parameter set, 1e5 draws)
    from ssmhip.kde import KDEBank, PointBatch

    sims = batched.BatchedSimulation("ddm", th, 100_000, seed=11, smooth_unif=True, launch=None)
    sims.run()
    bank = KDEBank(sims.rts, sims.choices, sims.job.offsets, [-1, 1])
    set_idx = np.concatenate([np.full(x_.shape[0], i) for i, x_ in enumerate(xs_parts)])
    pts, order = PointBatch.from_host(xs[:, 5], (xs[:, 6] > 0).astype(np.int64), k, 2, set_idx=set_idx)
    kll = bank.evaluate(pts).cpu().numpy()
    kde_at = np.empty_like(kll)
    kde_at[order] = kll
    kde_mae = float(np.mean(np.abs(kde_at - ref)))
    # roof: the surrogate's hidden unit u = 1 / (2^z + 1) (exp2 + rcp on the transcendental
    # unit) at full occupancy, over the network's 320 hidden units per row (100 + 100 + 120;
    # the forward computes 328, padding included), MFMA running beside them
    unit_rate = None if args.no_probe else roofline.probe_rate("hidden_unit", iters=2048)
    rec_roof = None if unit_rate is None else {
        "bound": "transcendental (hidden units)", "achieved": evals / 1e9, "peak": unit_rate / 320 / 1e9,
        "unit": "G rows/s", "frac": evals / (unit_rate / 320), "hidden_units_per_s": unit_rate}
    rec = {"inference_roofline": rec_roof, "label_sets": args.surrogate_sets, "label_samples_per_set": args.surrogate_samples,
           "label_s": label_s, "train_rows": int(train.shape[0]), "epochs": args.surrogate_epochs, "train_s": train_s,
           "train_rows_per_s": train.shape[0] * args.surrogate_epochs / train_s, "huber_history": hist[-3:],
           "val_mae_vs_kde_above_floor": net.val_mae, "val_mae_vs_kde_all": net.val_mae_all, "inference_rows": N, "ms_per_inference": ms, "evals_per_s": evals,
           "over_gpu_kde": evals / kde_evals_per_s,
           "mfma_tflops_algorithmic": flops / 1e12, "mfma_frac_of_bf16_peak": flops / 2.5e15,
           "mae_vs_analytic_theta0": mae, "kde_mae_vs_analytic_theta0": kde_mae, "analytic_rows": int(ref.size),
           "mae_vs_analytic_theta0_trained_on_1024x1e5_rows": mae_c4,
           "training_on_1024x1e5_rows": c4_fit}
    rec["inference_published_peak"] = {
        "mfma_bf16": {"unit": "TFLOP/s", "achieved": flops / 1e12, "peak": 2500.0, "frac": flops / 2.5e15,
                      "basis": "45,640 algorithmic FLOP per row; counters: MFMA busy 40.4% of SIMD cycles "
                               "(profiles/mlp_forward_4M_counters.md)"},
        "transcendental": {"unit": "T lane-instructions/s", "achieved": 658 * evals / 1e12,
                           "peak": PEAK_TRANS_PER_S / 1e12, "frac": 658 * evals / PEAK_TRANS_PER_S,
                           "basis": "658 v_exp_f32 + v_rcp_f32 per row (counters)"},
    }
    log(f"surrogate: {evals:.3e} evals/s ({rec['over_gpu_kde']:.0f}x GPU KDE), MAE vs analytic {mae:.3f} nats "
        f"(KDE labels: {kde_mae:.3f}; trained on the 1024 x 1e5 rows: {mae_c4})")
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-sets", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--launch", type=str, default="auto", help="'auto' or block,blocks_per_cu,refill_min")
    ap.add_argument("--no-probe", action="store_true", help="skip the VALU roof probe")
    ap.add_argument("--queue-order", choices=("expected", "index"), default="expected",
                    help="visiting order of the sets: longest expected walks first, or set index order")
    ap.add_argument("--headline-only", action="store_true", help="run the headline leg only")
    ap.add_argument("--legs", type=str, default="headline,ddm,accumulators,hssm,lan,surrogate",
                    help="comma list of legs to run (surrogate needs lan); without 'headline' the "
                         "JSON line carries no headline value (for profiling one leg)")
    ap.add_argument("--surrogate-epochs", type=int, default=150)
    ap.add_argument("--surrogate-sets", type=int, default=32768, help="parameter sets of KDE rows to train on")
    ap.add_argument("--surrogate-samples", type=int, default=1_000_000, help="simulations per set for the labels")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the workload shards by parameter set with no exchange, "
              "and no multi-GPU figure is produced in this build", file=sys.stderr)
        return 2

    import torch

    from ssmhip import batched, roofline

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    legs = {"headline"} if args.headline_only else set(args.legs.split(","))
    unknown = legs - {"headline", "ddm", "accumulators", "hssm", "lan", "surrogate"}
    if unknown or ("surrogate" in legs and "lan" not in legs):
        print(f"bad --legs {args.legs!r}: legs are headline, ddm, accumulators, hssm, lan, surrogate (surrogate needs "
              f"lan)", file=sys.stderr)
        return 2
    out = {"metric": "angle-DDM trajectories/s", "value": None, "unit": "trajectories/s", "n_gpus": 1,
           "legs": sorted(legs)}
    if "headline" in legs:
        out = headline_leg(args, torch, batched, roofline)
    if "ddm" in legs:
        out["ddm"] = ddm_leg(args)
    if "accumulators" in legs:
        out["accumulators"] = accumulator_leg(args)
    if "hssm" in legs:
        out["hssm"] = hssm_leg(args)
    if "lan" in legs:
        c4, lab, rows = lan_leg(args)
        out["lan_rows"] = c4
        if "surrogate" in legs:
            out["surrogate"] = surrogate_leg(args, lab, c4["kde_evals_per_s"], rows)
        del rows
    print(json.dumps(out), flush=True)
    return 0


def counter_value(key: str, profile: str = "em_angle_batched4096_counters.md"):
    """A figure of a committed counter summary (rocprofv3 --pmc runs of the same kernel):
    e.g. traffic_bytes_per_launch, valu_lane_instructions_per_step."""
    path = REPO / "profiles" / profile
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


# MI355X at its 2.4 GHz peak engine clock: 256 CUs x 4 SIMDs; a SIMD issues a wave64 FP32 FMA
# per quad-cycle (4 clocks), the rate of the published 157.3 TFLOP/s with packed FMAs
SIMD_QUAD_CYCLES_PER_S = 256 * 4 * 2.4e9 / 4
PEAK_FP32_FLOPS = 157.3e12
# transcendentals (v_exp/v_log/v_sin/v_sqrt/v_rcp_f32) occupy the VALU two quad-cycles
# (profiles/valu_issue_probes_counters.md): 1024 SIMDs x 64 lanes / 8 clocks x 2.4 GHz
PEAK_TRANS_PER_S = 256 * 4 * 64 / 8 * 2.4e9


def published_valu_view(steps_per_s: float):
    """The headline kernel against MI355X's published peaks, with the instruction counts per
    EM step from the counter summary of the same launch: VALU issue (wave-instructions per SIMD
    per quad-cycle, against the one per quad-cycle at which the FP32 peak is quoted), FP32
    arithmetic (3 FLOP per step) against 157.3 TFLOP/s, transcendentals against their
    two-quad-cycle ceiling."""
    ips = counter_value("valu_lane_instructions_per_step")
    tps = counter_value("trans_lane_instructions_per_step")
    if ips is None or tps is None:
        return None
    issue = ips / 64 * steps_per_s / SIMD_QUAD_CYCLES_PER_S
    return {
        "valu_issue": {"unit": "wave-instructions per SIMD per quad-cycle", "achieved": issue, "peak": 1.0,
                       "frac": issue,
                       "basis": f"{ips:.2f} lane-level VALU instructions per EM step (counters) x EM steps/s / "
                                "(1024 SIMDs x 2.4 GHz / 4); 1.0 = one FP32 FMA per quad-cycle, the rate of the "
                                "published 157.3 TFLOP/s; integer/bitwise streams issue faster "
                                "(profiles/valu_issue_probes_counters.md), so frac can exceed 1"},
        "fp32": {"unit": "TFLOP/s", "achieved": 3 * steps_per_s / 1e12, "peak": PEAK_FP32_FLOPS / 1e12,
                 "frac": 3 * steps_per_s / PEAK_FP32_FLOPS, "basis": "3 FLOP per EM step (the update)"},
        "transcendental": {"unit": "T lane-instructions/s", "achieved": tps * steps_per_s / 1e12,
                           "peak": PEAK_TRANS_PER_S / 1e12, "frac": tps * steps_per_s / PEAK_TRANS_PER_S,
                           "basis": f"{tps:.3f} transcendental lane-instructions per EM step (counters)"},
    }


def em_traffic_bytes():
    """HBM bytes per headline launch from the committed counter summary (two rocprofv3 --pmc runs)."""
    v = counter_value("traffic_bytes_per_launch")
    return None if v is None else int(v)


def headline_leg(args, torch, batched, roofline) -> dict:
    """Batched angle-DDM, 4096 ragged sets, one persistent hipGraph launch per step."""
    from ssmhip.em import LaunchParams

    theta, ns = batched.sample_angle_ragged(args.n_sets, seed=0)
    log(f"headline: {args.n_sets} angle sets, {int(ns.sum())} trajectories")
    launch = "auto" if args.launch == "auto" else LaunchParams(*[int(x) for x in args.launch.split(",")])
    # the headline batch tunes its launch shape on itself (~3 s), not on a sample
    sim = batched.BatchedSimulation("angle", theta, ns, delta_t=0.001, max_t=20.0, seed=0, launch=launch,
                                    tune_sample=None, queue_order=queue_order(args))
    lp = sim.launch_params
    log(f"autotuned launch: block={lp.block_size} blocks/CU={lp.blocks_per_cu} refill_min={lp.refill_min}")
    sim.capture()
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        sim.replay()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        sim.replay()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(end)
    ms = ms_total / args.steps
    steps_per_launch = sim.steps()
    traj = sim.n_total
    value = traj / (ms / 1e3)
    steps_per_s = steps_per_launch / (ms / 1e3)
    # roof: the same Philox + Box–Muller + angle-step loop, branch-free, at full occupancy
    ideal = None if args.no_probe else roofline.probe_rate("em_block", iters=2048)
    slots = sim.lane_slots()
    util = steps_per_launch / slots if slots else None
    log(f"{ms:.3f} ms/launch, {value:.4e} traj/s, {steps_per_launch / (ms / 1e3):.4e} EM steps/s "
        f"(wall {wall:.2f} s, lane utilisation {util})")

    # hand-over to the host (what a caller of the array-returning drop-ins also pays): the
    # device-to-host copy of rts + choices into pinned buffers, timed after the launches
    h_rts = torch.empty(sim.rts.shape, dtype=sim.rts.dtype, pin_memory=True)
    h_ch = torch.empty(sim.choices.shape, dtype=sim.choices.dtype, pin_memory=True)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    h_rts.copy_(sim.rts, non_blocking=True)
    h_ch.copy_(sim.choices, non_blocking=True)
    c1.record(stream)
    c1.synchronize()
    d2h_ms = c0.elapsed_time(c1)
    d2h_bytes = h_rts.numel() * h_rts.element_size() + h_ch.numel() * h_ch.element_size()
    # sanity: choices are +-1 and rts finite or the omission sentinel
    ch = h_ch.numpy()
    rts = h_rts.numpy()
    assert np.isin(ch, (-1, 0, 1)).all() and np.isfinite(rts).all()

    out = {
        "metric": "angle-DDM trajectories/s",
        "value": value,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched angle-DDM, 4096 sets, theta~U(angle bounds), n_i=floor(10^U(3,5)), dt=1e-3, "
                        "max_t=20, one persistent hipGraph launch",
            "n_sets": int(args.n_sets),
            "trajectories": int(traj),
            "em_steps_per_launch": int(steps_per_launch),
            "em_steps_per_s": steps_per_launch / (ms / 1e3),
            "launch": {"block": lp.block_size, "blocks_per_cu": lp.blocks_per_cu, "refill_min": lp.refill_min,
                       "queue_block": lp.queue_block},
            "queue_order": args.queue_order,
            "lane_utilisation": util,
            "host_handover": {"d2h_ms": d2h_ms, "bytes": int(d2h_bytes), "GB_per_s": d2h_bytes / d2h_ms / 1e6,
                              "trajectories_per_s_incl_d2h": traj / ((ms + d2h_ms) / 1e3)},
        },
        "roofline": {
            "bound": "valu",
            "achieved": steps_per_s / 1e9,
            "peak": None if ideal is None else ideal / 1e9,
            "unit": "G EM-steps/s",
            "frac": None if ideal is None else steps_per_s / ideal,
            "traffic": em_traffic_bytes(),
            "traffic_basis": "HBM bytes per launch: rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (factors from a "
                             "1 GiB copy in the same runs), profiles/em_angle_batched4096_counters.md; algorithmic "
                             f"bytes {int(traj) * 5 + int(args.n_sets) * 264}",
            "peak_basis": "measured: branch-free Philox4x32-10 + Box-Muller + angle-step loop at full occupancy "
                          "(VALU issue bound; no MFMA, ~5 B of HBM traffic per trajectory)",
            "published_peak": published_valu_view(steps_per_s),
        },
    }
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = 2 * n * 4
    if not args.no_cpu:
        cb = cpu_baseline_leg(theta, ns, args.cpu_seconds)
        cb.pop("calibration_rate", None)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    sim.close()
    return out


if __name__ == "__main__":
    sys.exit(main())
