"""GPU parity of the EM kernels against the oracle (run on an MI355X).

Replay mode: the kernel and the float32 NumPy transcription consume the same
normal-increment buffer; every trajectory must give the identical choice and
an rt within one dt (in practice bit-identical). Native mode: Philox +
Box–Muller on the GPU against the C++ MT19937 + Ziggurat restatement and the
Wiener closed form, at 1e5 samples: KS p > 0.01 per choice, choice
probabilities and the 0.1..0.9 rt quantiles within 3 Monte-Carlo SEs.
"""

import math

import numpy as np
import pytest
from scipy import stats

from em_cases import compare_samples as _compare_samples
from em_cases import quantile_check as _quantile_check
from oracle import cpu_baseline, ddm_analytic
from oracle import em_numpy as E
from oracle import philox as PH

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = 0.001


def _ssm():
    from ssmhip import _lib as L
    from ssmhip import em

    return L, em


# ----------------------------------------------------------------------------- RNG
def test_philox_raw_bits_match_oracle(torch_gpu):
    import ctypes as C

    from ssmhip import _device as D
    from ssmhip import _lib as L

    torch = torch_gpu
    n = 4096
    raw = D.empty((n, 4), torch.int32)
    nz = D.empty((n, 4), torch.float32)
    seed = 0x0123456789ABCDEF
    rc = L.load().ssm_philox_fill(seed, 5, 7, 11, 13, n, D.ptr(raw), D.ptr(nz), D.stream_ptr())
    L.check(rc)
    got = raw.cpu().numpy().view(np.uint32)
    ctr = np.zeros((n, 4), np.uint32)
    ctr[:, 0] = 5 + np.arange(n, dtype=np.uint32)
    ctr[:, 1], ctr[:, 2], ctr[:, 3] = 7, 11, 13
    expect = PH.philox4x32_10(ctr, PH.key_from_seed(seed))
    assert np.array_equal(got, expect)
    # Box–Muller with hardware log2/sin/cos vs float64
    ref = PH.normals4(ctr, seed)
    err = np.abs(nz.cpu().numpy().astype(np.float64) - ref)
    assert err.max() < 5e-5 * max(1.0, np.abs(ref).max())
    assert stats.kstest(nz.cpu().numpy().reshape(-1), "norm").pvalue > 1e-3


# ----------------------------------------------------------------------------- replay parity
def _ddm_case(bname, K, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-3, 3, K).astype(F32)
    a = rng.uniform(0.3, 2.5, K).astype(F32)
    z = rng.uniform(0.1, 0.9, K).astype(F32)
    t = rng.uniform(0.0, 2.0, K).astype(F32)
    bp = {"a": a, "theta": rng.uniform(-0.1, 1.3, K).astype(F32), "alpha": rng.uniform(0.3, 5, K).astype(F32),
          "beta": rng.uniform(0.3, 7, K).astype(F32)}
    bp = {k: bp[k] for k in E.BOUNDARIES[bname][1]}
    return v, a, z, t, bp


def _run_gpu_ddm(v, a, z, t, bp, bname, N, normals, unif, smooth, use_table):
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    K = v.shape[0]
    s = np.ones(K, F32)
    fun = {"constant": BF.constant, "angle": BF.angle, "weibull_cdf": BF.weibull_cdf}[bname]
    code = L.BOUNDARY_TABLE if use_table else BF.kernel_boundary_code(fun)
    rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), s, 20.0, DT, code, bp)
    table = em.boundary_table(fun, bp, K, 20.0, DT) if use_table else None
    job = em.make_job(L.MODEL_DDM, code, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0, seed=1,
                      smooth_unif=smooth, table=table, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    return job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy(), job.steps()


@pytest.mark.parametrize("bname,use_table", [("constant", False), ("angle", False), ("weibull_cdf", True),
                                             ("angle", True)])
def test_ddm_replay_matches_oracle_trajectory_by_trajectory(torch_gpu, bname, use_table):
    K, N, L_ = 16, 500, 1031
    print(f"[case] ddm replay {bname} table={use_table}", flush=True)
    v, a, z, t, bp = _ddm_case(bname, K, 21)
    rng = np.random.default_rng(3)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 boundary=bname, bparams=bp, smooth_unif=True, normals=normals, uniforms=unif)
    r1, c1, steps = _run_gpu_ddm(v, a, z, t, bp, bname, N, normals, unif, True, use_table)
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert steps == int(s0.sum())


@pytest.mark.parametrize("dt,max_t", [(0.000731, 3.0), (0.0025, 10.0), (0.01, 25.0)])
def test_ddm_replay_times_at_other_step_sizes(torch_gpu, dt, max_t):
    """Walks far from their deadlines take their time from the library's table of t after k
    steps (k float32 additions of dt from 0). At step sizes other than 1e-3, with walks long
    enough that t crosses several powers of two, every rt must still equal the oracle's
    sequentially accumulated one."""
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    K, N, L_ = 12, 200, 1031
    print(f"[case] ddm replay at dt {dt}, max_t {max_t}", flush=True)
    rng = np.random.default_rng(61)
    v = rng.uniform(-0.4, 0.4, K).astype(F32)
    a = rng.uniform(2.0, 3.0, K).astype(F32)
    z = rng.uniform(0.3, 0.7, K).astype(F32)
    t = rng.uniform(0.0, 0.5, K).astype(F32)
    dt32 = float(F32(dt))
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    bp = {"a": a, "theta": rng.uniform(-0.1, 0.1, K).astype(F32)}
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, max_t=max_t,
                                 delta_t=dt32, boundary="angle", bparams=bp, smooth_unif=True, normals=normals,
                                 uniforms=unif)
    rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), max_t, dt32, L.BOUNDARY_ANGLE, bp)
    job = em.make_job(L.MODEL_DDM, L.BOUNDARY_ANGLE, 1, rec, em.uniform_offsets(K, N), delta_t=dt32, max_t=max_t,
                      seed=1, smooth_unif=True, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    r1, c1 = job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy()
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert job.steps() == int(s0.sum())
    # walks long enough for t to pass 1, 2 and 4 (t's float32 spacing changes at each)
    assert np.max(s0) * dt32 > 4.0 or np.max(s0) >= BF.n_steps(max_t, dt32) - 1


@pytest.mark.parametrize("bname", ["constant", "angle"])
def test_ddm_replay_with_binding_deadlines_and_a_short_time_grid(torch_gpu, bname):
    """Walks that end by their deadline (t > deadline - t0) or by reaching the end of t_s,
    both a few steps into a block or an iteration: the kernel's iterations that skip the index
    and deadline tests must hand over to the tested steps exactly where those tests bind."""
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    K, N, L_ = 24, 400, 1031
    print(f"[case] ddm replay deadlines + short t_s {bname}", flush=True)
    rng = np.random.default_rng(44)
    v = rng.uniform(-0.3, 0.3, K).astype(F32)        # slow drift: many walks reach the cut-offs
    a = rng.uniform(1.5, 3.0, K).astype(F32)         # wide bounds
    z = rng.uniform(0.3, 0.7, K).astype(F32)
    t = rng.uniform(0.0, 0.5, K).astype(F32)
    dl = (t + rng.uniform(0.005, 0.4, K)).astype(F32)  # deadlines 5 to 400 steps after t
    dl[::4] = F32(999.0)                              # some sets run into the end of t_s instead
    bp = {"a": a, "theta": rng.uniform(-0.1, 0.2, K).astype(F32)}
    bp = {k: bp[k] for k in E.BOUNDARIES[bname][1]}
    max_t = 0.35
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, dl, np.ones(K, F32), n_samples=N, max_t=max_t, delta_t=DT,
                                 boundary=bname, bparams=bp, smooth_unif=True, normals=normals, uniforms=unif)
    fun = {"constant": BF.constant, "angle": BF.angle}[bname]
    code = BF.kernel_boundary_code(fun)
    rec = em.ddm_records(v, a, z, t, dl, np.ones(K, F32), max_t, DT, code, bp)
    job = em.make_job(L.MODEL_DDM, code, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=max_t, seed=1,
                      smooth_unif=True, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    r1, c1 = job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy()
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert job.steps() == int(s0.sum())
    # the cut-offs bind: walks omitted by their deadline, and walks that ran to the end of t_s
    n_steps = int(round(max_t / DT)) + 1
    assert np.sum(r1 == -999.0) > 0.05 * r1.size
    assert np.any(s0 >= n_steps - 1)


def test_ddm_start_point_rounding_matches_oracle(torch_gpu):
    # Sets whose walk length depends on the last bit of the start point: v = 0, s = 0 keeps y at
    # its start and the angle bound falls by half an ulp per step (tests/em_cases.py). The DDM
    # kernel must end where the float start of ddm_models.pyx:709 puts the walk, and the full
    # DDM kernel (zero variability) where the double-literal start of :2305 does.
    import torch

    from em_cases import start_sensitive_ddm_cases

    L, em = _ssm()
    v, a, z, th, _ = start_sensitive_ddm_cases(2000)
    K = v.shape[0]
    print(f"[case] start-point rounding, {K} sets", flush=True)
    assert K >= 50
    zeros, dl = np.zeros(K, F32), np.full(K, 999, F32)
    bp = {"a": a, "theta": th}
    normals = np.zeros((K, 4), F32)
    common = dict(n_samples=1, delta_t=DT, boundary="angle", bparams=bp, normals=normals)
    var = {"sz": zeros, "sv": zeros, "st": zeros}
    unif = np.full((K, 3), 0.5, F32)
    for model, extra, ru in ((L.MODEL_DDM, {}, None), (L.MODEL_FULL_DDM, var, unif)):
        r0, c0, s0 = E.ddm_flexbound(v, a, z, zeros, dl, zeros, uniforms=ru, **extra, **common)
        rec = em.ddm_records(v, a, z, zeros, dl, zeros, 20.0, DT, L.BOUNDARY_ANGLE, bp, **extra)
        job = em.make_job(model, L.BOUNDARY_ANGLE, 1, rec, em.uniform_offsets(K, 1), delta_t=DT, max_t=20.0, seed=1,
                          smooth_unif=False, replay_normals=normals, replay_uniforms=ru)
        job.launch()
        torch.cuda.synchronize()
        assert np.array_equal(job.choices[:K].cpu().numpy(), c0)
        assert np.array_equal(job.rts[:K].cpu().numpy(), r0)
        assert job.steps() == int(s0.sum())


def test_ddm_weibull_in_register_vs_table(torch_gpu):
    # analytic Weibull uses hardware exp2/log2, the table uses NumPy float32 exp/power:
    # boundary values differ by ulps, so rare crossings may shift; bound the rate.
    K, N, L_ = 16, 1000, 1031
    v, a, z, t, bp = _ddm_case("weibull_cdf", K, 22)
    rng = np.random.default_rng(4)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, _ = _run_gpu_ddm(v, a, z, t, bp, "weibull_cdf", N, normals, unif, False, True)
    r1, c1, _ = _run_gpu_ddm(v, a, z, t, bp, "weibull_cdf", N, normals, unif, False, False)
    assert np.mean(c0 != c1) < 2e-3
    assert np.mean(np.abs(r0 - r1) > DT) < 2e-3


@pytest.mark.parametrize("model", ["race", "lca"])
@pytest.mark.parametrize("bname", ["constant", "weibull_cdf", "angle"])
def test_accumulator_replay_matches_oracle(torch_gpu, model, bname):
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    print(f"[case] {model} replay {bname}", flush=True)
    K, P, N, L_ = 8, 4, 400, 2053
    rng = np.random.default_rng(31)
    v = rng.uniform(0, 2.5, (K, P)).astype(F32)
    z = rng.uniform(0, 0.5, (K, P)).astype(F32)
    a = rng.uniform(1, 3, K).astype(F32)
    t = rng.uniform(0, 2, K).astype(F32)
    g = rng.uniform(-1, 1, K).astype(F32)
    b = rng.uniform(-1, 1, K).astype(F32)
    bp = {"a": a, "alpha": rng.uniform(0.3, 5, K).astype(F32), "beta": rng.uniform(0.3, 7, K).astype(F32),
          "theta": rng.uniform(-0.1, 1.45, K).astype(F32)}
    bp = {k: bp[k] for k in E.BOUNDARIES[bname][1]}
    kw = dict(g=g, b=b) if model == "lca" else {}
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.accumulators(model, v, z, t, np.ones((K, 1), F32), np.full(K, 999, F32), n_samples=N,
                                delta_t=DT, boundary=bname, bparams=bp, smooth_unif=True, normals=normals,
                                uniforms=unif, **kw)
    fun = {"constant": BF.constant, "weibull_cdf": BF.weibull_cdf, "angle": BF.angle}[bname]
    # constant and Weibull read the boundary table; the angle bound is evaluated in registers
    code = L.BOUNDARY_ANGLE if bname == "angle" else L.BOUNDARY_TABLE
    rec = em.accumulator_records(v, z, t, np.ones((K, 1), F32), np.full(K, 999, F32), 20.0, DT, code, bp, **kw)
    table = em.boundary_table(fun, bp, K, 20.0, DT) if code == L.BOUNDARY_TABLE else None
    mdl = L.MODEL_RACE if model == "race" else L.MODEL_LCA
    job = em.make_job(mdl, code, P, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0, seed=1,
                      smooth_unif=True, table=table, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    r1 = job.rts[: K * N].cpu().numpy()
    c1 = job.choices[: K * N].cpu().numpy()
    assert np.array_equal(c0, c1)
    assert np.all(np.abs(r0 - r1) <= DT)
    assert np.mean(r0 == r1) > 0.999
    assert job.steps() == int(s0.sum())


def test_replay_is_independent_of_launch_shape(torch_gpu):
    import torch

    L, em = _ssm()
    K, N, L_ = 8, 700, 257
    v, a, z, t, bp = _ddm_case("angle", K, 5)
    rng = np.random.default_rng(6)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), 20.0, DT, L.BOUNDARY_ANGLE, bp)
    outs = []
    for lp in [em.LaunchParams(64, 1, 1, 1), em.LaunchParams(256, 8, 16, 64), em.LaunchParams(1024, 2, 64, 4096)]:
        job = em.make_job(L.MODEL_DDM, L.BOUNDARY_ANGLE, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0,
                          seed=1, smooth_unif=False, replay_normals=normals)
        job.launch(lp)
        torch.cuda.synchronize()
        outs.append((job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy()))
    for r, c in outs[1:]:
        assert np.array_equal(r, outs[0][0]) and np.array_equal(c, outs[0][1])


@pytest.mark.parametrize("model", ["ddm", "lca"])
def test_native_output_is_independent_of_launch_shape_on_ragged_sets(torch_gpu, model):
    """Counter-based normals: the same seed gives the same trajectories whichever wave, lane,
    local queue range or refill point a trajectory lands on; ragged and empty sets included."""
    import torch

    L, em = _ssm()
    sizes = np.array([0, 1, 3, 700, 0, 64, 65, 1000, 2, 129], np.int64)
    K = sizes.size
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    rng = np.random.default_rng(11)
    if model == "ddm":
        v, a, z, t, bp = _ddm_case("angle", K, 12)
        rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), 20.0, DT, L.BOUNDARY_ANGLE, bp)
        kind, bound, P = L.MODEL_DDM, L.BOUNDARY_ANGLE, 1
    else:
        P = 3
        v = rng.uniform(0.2, 1.5, (K, P)).astype(F32)
        z = rng.uniform(0.0, 0.3, (K, P)).astype(F32)
        rec = em.accumulator_records(v, z, rng.uniform(0.1, 0.5, K).astype(F32), np.ones((K, P), F32),
                                     np.full(K, 999, F32), 20.0, DT, L.BOUNDARY_CONSTANT,
                                     {"a": rng.uniform(1.0, 2.5, K).astype(F32)},
                                     g=rng.uniform(0, 0.3, K).astype(F32), b=rng.uniform(0, 0.3, K).astype(F32))
        kind, bound = L.MODEL_LCA, L.BOUNDARY_CONSTANT
    outs = []
    expected = np.argsort(-em.expected_steps(rec, kind, P, 20.0, DT), kind="stable")
    cases = [(em.LaunchParams(64, 1, 1, 1), None), (em.LaunchParams(64, 1, 1, 64), None),
             (em.LaunchParams(256, 8, 8, 64), None), (em.LaunchParams(512, 4, 3, 7), None),
             (em.LaunchParams(1024, 2, 64, 4096), None),
             # the queue's visiting order of the sets changes nothing either
             (em.LaunchParams(256, 8, 8, 64), expected), (em.LaunchParams(64, 1, 1, 1), rng.permutation(K)),
             (em.LaunchParams(512, 4, 3, 7), np.arange(K)[::-1])]
    for lp, order in cases:
        job = em.make_job(kind, bound, P, rec, offs, delta_t=DT, max_t=20.0, seed=5, smooth_unif=True)
        job.set_queue_order(order)
        job.launch(lp)
        torch.cuda.synchronize()
        n = int(offs[-1])
        outs.append((job.rts[:n].cpu().numpy(), job.choices[:n].cpu().numpy(), job.steps()))
    for r, c, st in outs[1:]:
        assert np.array_equal(r, outs[0][0]) and np.array_equal(c, outs[0][1]) and st == outs[0][2]
    assert np.all(outs[0][0] > 0)


@pytest.mark.parametrize("kind", ["ddm", "ornstein"])
def test_small_job_in_four_walk_waves_equals_the_same_walks_in_full_waves(torch_gpu, kind):
    """A job of at most 16 walks per SIMD runs its native Gaussian walks 4 to a wave (each
    stepping in lane-shift chains from its first step), up to 40 walks per SIMD 16 to a wave;
    a larger job fills every lane. The same 10,000-walk set alone, as the first set of a
    25,000-walk job and of a 60,000-walk job must give identical rts and choices."""
    import torch

    L, em = _ssm()
    cus = int(L.load().ssm_device_cu_count())
    n0 = 10_000
    assert n0 <= 64 * cus < 25_000 <= 160 * cus < 60_000  # 4-walk, 16-walk and full waves
    v, a, z, t = (np.array(x, F32) for x in ([0.8, -0.5], [1.2, 1.6], [0.45, 0.55], [0.2, 0.3]))
    bp = {"a": a, "theta": np.array([0.25, 0.1], F32)}
    model = L.MODEL_DDM if kind == "ddm" else L.MODEL_ORNSTEIN
    extra = {} if kind == "ddm" else {"g": np.array([0.3, -0.2], F32)}
    rec = em.ddm_records(v, a, z, t, np.full(2, 999, F32), np.ones(2, F32), 20.0, DT, L.BOUNDARY_ANGLE, bp, **extra)
    print(f"[case] small job in 4-walk waves, {kind}", flush=True)
    alone = em.make_job(model, L.BOUNDARY_ANGLE, 1, rec[:1], np.array([0, n0], np.int64), delta_t=DT, max_t=20.0,
                        seed=3, smooth_unif=True)
    alone.launch()
    torch.cuda.synchronize()
    for total in (25_000, 60_000):
        both = em.make_job(model, L.BOUNDARY_ANGLE, 1, rec, np.array([0, n0, total], np.int64), delta_t=DT,
                           max_t=20.0, seed=3, smooth_unif=True)
        both.launch()
        torch.cuda.synchronize()
        assert np.array_equal(alone.rts[:n0].cpu().numpy(), both.rts[:n0].cpu().numpy())
        assert np.array_equal(alone.choices[:n0].cpu().numpy(), both.choices[:n0].cpu().numpy())
    assert alone.steps() > 100 * n0  # walks of hundreds of steps


@pytest.mark.parametrize("model", ["ddm", "ddm_deadline", "ddm_table", "ddm_legacy", "ornstein", "full_ddm"])
def test_native_tail_walks_equal_full_wave_walks(torch_gpu, model):
    """Once at most 32 walks are left in a wave, the wave's idle lanes draw the Philox blocks
    of the walking lanes, and from 8 walks on each walk forms its positions in a group of lanes.
    A set of 64 walks launched alone ends that way: its wave holds all 64 and runs its last 32
    in those modes, down to one walk. The same set launched ahead of 100 sets of slow walks,
    with waves that refill as soon as a lane is free, runs every walk of it in a full wave.
    Both launches must give identical rts and choices."""
    import torch

    L, em = _ssm()
    K = 101
    rng = np.random.default_rng(23)
    v = np.concatenate([[0.7], np.zeros(K - 1)]).astype(F32)  # set 0: the one under test
    a = np.concatenate([[1.4], np.full(K - 1, 2.5)]).astype(F32)  # slow sets: E[T] = 6.25 s
    z = np.concatenate([[0.45], np.full(K - 1, 0.5)]).astype(F32)
    t = np.full(K, 0.1, F32)
    dl, s = np.full(K, 999, F32), np.ones(K, F32)
    table = None
    if model in ("ddm", "ddm_deadline"):
        if model == "ddm_deadline":  # the longest walks end at the deadline, inside a tail
            dl[0] = 1.9
        bp = {"a": a, "theta": np.concatenate([[0.3], np.zeros(K - 1)]).astype(F32)}
        kind, bound = L.MODEL_DDM, L.BOUNDARY_ANGLE
        rec = em.ddm_records(v, a, z, t, dl, s, 20.0, DT, bound, bp)
    elif model == "ddm_table":  # a collapsing Weibull bound read from the per-set table
        from ssmhip import boundary_functions as BF

        bp = {"a": a, "alpha": np.full(K, 2.0, F32), "beta": np.full(K, 4.0, F32)}
        kind, bound = L.MODEL_DDM, L.BOUNDARY_TABLE
        rec = em.ddm_records(v, a, z, t, dl, s, 20.0, DT, bound, bp)
        table = em.boundary_table(BF.weibull_cdf, bp, K, 20.0, DT)
    elif model == "ddm_legacy":  # bounds [0, a], start z a
        kind, bound = L.MODEL_DDM_LEGACY, L.BOUNDARY_CONSTANT
        rec = em.ddm_records(v, a, z, t, dl, s, 20.0, DT, bound, {"a": a})
    elif model == "ornstein":
        kind, bound = L.MODEL_ORNSTEIN, L.BOUNDARY_CONSTANT
        rec = em.ddm_records(v, a, z, t, dl, s, 20.0, DT, bound, {"a": a}, g=np.full(K, 0.4, F32))
    else:
        kind, bound = L.MODEL_FULL_DDM, L.BOUNDARY_CONSTANT
        rec = em.ddm_records(v, a, z, t, dl, s, 20.0, DT, bound, {"a": a}, sz=np.full(K, 0.1, F32),
                             sv=np.full(K, 0.5, F32), st=np.full(K, 0.05, F32))
    print(f"[case] tail walks {model}", flush=True)
    n0 = 64
    alone = em.make_job(kind, bound, 1, rec[:1], np.array([0, n0], np.int64), delta_t=DT, max_t=20.0, seed=9,
                        smooth_unif=True, table=None if table is None else table[:1])
    alone.launch(em.LaunchParams(256, 8, 8, 64))
    torch.cuda.synchronize()
    sizes = np.concatenate([[n0], np.full(K - 1, 1000)]).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    full = em.make_job(kind, bound, 1, rec, offs, delta_t=DT, max_t=20.0, seed=9, smooth_unif=True, table=table)
    full.launch(em.LaunchParams(64, 1, 1, 1))
    torch.cuda.synchronize()
    r_alone, c_alone = alone.rts[:n0].cpu().numpy(), alone.choices[:n0].cpu().numpy()
    assert np.array_equal(r_alone, full.rts[:n0].cpu().numpy())
    assert np.array_equal(c_alone, full.choices[:n0].cpu().numpy())
    # the set's walks span a range of lengths, so the alone launch's wave shrank through every
    # tail regime (32, 16, 8, 4 and fewer walks)
    if model == "ddm_deadline":
        assert np.sum(r_alone == -999) >= 2  # walks cut at the deadline
        r_alone = r_alone[r_alone > 0]
    steps = np.round((r_alone - 0.1) / DT)
    assert np.ptp(steps) > 1000, (steps.min(), steps.max())


def test_replay_is_independent_of_queue_order_on_ragged_sets(torch_gpu):
    """Replay rows are addressed by the trajectory's set-major index, whatever queue position
    it is drawn from: the oracle's trajectories come back in any visiting order of the sets."""
    import torch

    L, em = _ssm()
    sizes = np.array([5, 0, 300, 1, 77, 0, 129], np.int64)
    K, L_ = sizes.size, 301
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs[-1])
    print(f"[case] replay queue order, sizes {sizes.tolist()}", flush=True)
    v, a, z, t, bp = _ddm_case("angle", K, 21)
    rng = np.random.default_rng(22)
    normals = rng.standard_normal((n, L_)).astype(F32)
    uniforms = rng.uniform(size=n).astype(F32)
    rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), 20.0, DT, L.BOUNDARY_ANGLE, bp)
    outs = []
    for order in (None, rng.permutation(K), np.arange(K)[::-1]):
        job = em.make_job(L.MODEL_DDM, L.BOUNDARY_ANGLE, 1, rec, offs, delta_t=DT, max_t=20.0, seed=1,
                          smooth_unif=True, replay_normals=normals, replay_uniforms=uniforms)
        job.set_queue_order(order)
        job.launch(em.LaunchParams(64, 1, 1, 3))
        torch.cuda.synchronize()
        outs.append((job.rts[:n].cpu().numpy(), job.choices[:n].cpu().numpy()))
    for r, c in outs[1:]:
        assert np.array_equal(r, outs[0][0]) and np.array_equal(c, outs[0][1])
    for k in range(K):  # each set against the oracle on its own rows
        lo, hi = offs[k], offs[k + 1]
        if hi == lo:
            continue
        one = lambda x: x[k:k + 1]
        r0, c0, _ = E.ddm_flexbound(one(v), one(a), one(z), one(t), np.full(1, 999, F32), np.ones(1, F32),
                                    n_samples=int(hi - lo), delta_t=DT, boundary="angle",
                                    bparams={q: one(w) for q, w in bp.items()}, smooth_unif=True,
                                    normals=normals[lo:hi], uniforms=uniforms[lo:hi])
        assert np.array_equal(c0, outs[1][1][lo:hi])
        assert np.all(np.abs(r0 - outs[1][0][lo:hi]) <= DT)


# ----------------------------------------------------------------------------- native RNG statistics
def test_ddm_native_single_set_1e5_matches_cpu_restatement_and_closed_form(torch_gpu):
    import torch

    L, em = _ssm()
    N = 100_000
    v, a, z, t = F32(1.0), F32(1.5), F32(0.5), F32(0.2)
    rec = em.ddm_records([v], [a], [z], [t], [999.0], [1.0], 20.0, DT, L.BOUNDARY_CONSTANT, {"a": [a]})
    job = em.make_job(L.MODEL_DDM, L.BOUNDARY_CONSTANT, 1, rec, em.uniform_offsets(1, N), delta_t=DT, max_t=20.0,
                      seed=0, smooth_unif=False)
    job.launch()
    torch.cuda.synchronize()
    rg, cg = job.rts[:N].cpu().numpy(), job.choices[:N].cpu().numpy()
    table = E.boundary_table("constant", {"a": [a]}, 1, 20.0, DT)
    rc_, cc_, _ = cpu_baseline.em_simulate(0, [v], [z], [E.sqrt_dt32(DT)], [t], [999.0], [20.0], table,
                                           np.array([0, N], np.int64), dt=DT, seed=12345)
    _compare_samples(rg, cg, rc_, cc_, (-1, 1))
    a_eff = float(a) + ddm_analytic.SIEGMUND_BETA * math.sqrt(DT)
    p = ddm_analytic.p_upper(1.0, a_eff, 0.5)
    assert abs(np.mean(cg == 1) - p) < 3 * math.sqrt(p * (1 - p) / N)
    # density against Navarro–Fuss at the Siegmund-corrected bound: histogram chi-square
    for c in (-1, 1):
        x = rg[cg == c] - 0.2
        edges = np.quantile(x, np.linspace(0, 1, 21))
        pdf_mass = []
        for lo, hi in zip(edges[:-1], edges[1:]):
            grid = np.linspace(lo, hi, 64)
            pdf_mass.append(np.trapz(ddm_analytic.density(grid, 1.0, a_eff, 0.5, c), grid))
        pdf_mass = np.array(pdf_mass) / np.sum(pdf_mass)
        obs = np.histogram(x, edges)[0]
        exp_ = pdf_mass * obs.sum()
        chi2 = np.sum((obs - exp_) ** 2 / exp_)
        assert stats.chi2.sf(chi2, df=19) > 1e-3, (c, chi2)


@pytest.mark.parametrize("s", [0.5, 2.0])
def test_ddm_native_noise_scale_matches_cpu_restatement_and_closed_form(torch_gpu, s):
    """Native walks draw their normals already scaled by sqrt(dt) s (the Box–Muller radius
    carries s^2). At s != 1 they must match the C++/OpenMP restatement, which multiplies
    sqrt(dt) s by a unit normal, and the Wiener closed forms of the rescaled problem
    (v / s, a / s, unit noise), with the Siegmund correction beta sqrt(dt) on a / s."""
    import torch

    L, em = _ssm()
    N = 100_000
    # z = 1/2: the start stays midway between the Siegmund-corrected bounds
    v, a, z, t = F32(0.5), F32(1.0), F32(0.5), F32(0.2)
    rec = em.ddm_records([v], [a], [z], [t], [999.0], [s], 20.0, DT, L.BOUNDARY_CONSTANT, {"a": [a]})
    job = em.make_job(L.MODEL_DDM, L.BOUNDARY_CONSTANT, 1, rec, em.uniform_offsets(1, N), delta_t=DT, max_t=20.0,
                      seed=3, smooth_unif=False)
    job.launch()
    torch.cuda.synchronize()
    rg, cg = job.rts[:N].cpu().numpy(), job.choices[:N].cpu().numpy()
    table = E.boundary_table("constant", {"a": [a]}, 1, 20.0, DT)
    sqrt_st = F32(E.sqrt_dt32(DT) * F32(s))
    rc_, cc_, _ = cpu_baseline.em_simulate(0, [v], [z], [sqrt_st], [t], [999.0], [20.0], table,
                                           np.array([0, N], np.int64), dt=DT, seed=777)
    _compare_samples(rg, cg, rc_, cc_, (-1, 1))
    a_eff = float(a) / s + ddm_analytic.SIEGMUND_BETA * math.sqrt(DT)
    p = ddm_analytic.p_upper(float(v) / s, a_eff, float(z))
    assert abs(np.mean(cg == 1) - p) < 3 * math.sqrt(p * (1 - p) / N)
    dec = rg.astype(np.float64) - float(t)
    m = ddm_analytic.mean_decision_time(float(v) / s, a_eff, float(z))
    assert abs(dec.mean() - m) < 4 * dec.std() / math.sqrt(N) + 2 * DT, (dec.mean(), m)


@pytest.mark.parametrize("bname", ["angle", "weibull_cdf"])
def test_ddm_native_in_register_bounds_match_cpu_restatement(torch_gpu, bname):
    """Native Philox walks with the angle or Weibull bound evaluated in registers, against the
    C++/OpenMP restatement of the parallel branch (MT19937 + Ziggurat walking a NumPy boundary
    table) at 1e5 samples per set. Set 0 has opening bounds (theta < 0) or a slow collapse and a
    small drift, so a share of its walks runs to max_t."""
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    N, K = 100_000, 3
    v = np.array([0.1, 1.5, -0.8], F32)
    z = np.array([0.5, 0.4, 0.6], F32)
    t = np.array([0.3, 0.1, 0.5], F32)
    if bname == "angle":
        a = np.array([2.0, 1.2, 1.5], F32)
        bp = {"a": a, "theta": np.array([-0.1, 1.2, 0.4], F32)}
    else:
        a = np.array([5.0, 1.2, 1.5], F32)
        bp = {"a": a, "alpha": np.array([0.5, 3.0, 1.5], F32), "beta": np.array([200.0, 1.0, 3.0], F32)}
    print(f"[case] ddm native {bname} vs cpu restatement", flush=True)
    fun = {"angle": BF.angle, "weibull_cdf": BF.weibull_cdf}[bname]
    code = BF.kernel_boundary_code(fun)
    assert code != L.BOUNDARY_TABLE
    dl = np.full(K, 999.0, F32)
    rec = em.ddm_records(v, a, z, t, dl, np.ones(K, F32), 20.0, DT, code, bp)
    job = em.make_job(L.MODEL_DDM, code, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0, seed=4,
                      smooth_unif=False)
    job.launch()
    torch.cuda.synchronize()
    rg, cg = job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy()
    table = E.boundary_table(bname, bp, K, 20.0, DT)
    dl_tmp = np.minimum(F32(20.0), (dl - t).astype(F32)).astype(F32)
    rc_, cc_, _ = cpu_baseline.em_simulate(0, v, z, np.full(K, E.sqrt_dt32(DT)), t, dl, dl_tmp, table,
                                           np.arange(K + 1, dtype=np.int64) * N, dt=DT, seed=31)
    assert np.mean(rc_[:N] > 20.0) > 0.01 and np.mean(rg[:N] > 20.0) > 0.01  # walks that reach max_t
    for k in range(K):
        s_ = slice(k * N, (k + 1) * N)
        _compare_samples(rg[s_], cg[s_], rc_[s_], cc_[s_], (-1, 1))


@pytest.mark.parametrize("model", ["race", "lca"])
def test_accumulators_native_match_cpu_restatement(torch_gpu, model):
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    N, P = 100_000, 4
    v = np.array([[1.2, 0.8, 0.5, 1.0]], F32)
    z = np.array([[0.2, 0.1, 0.3, 0.25]], F32)
    a = np.array([2.0], F32)
    t = np.array([0.3], F32)
    g = np.array([0.4], F32)
    b = np.array([0.3], F32)
    bp = {"a": a, "alpha": np.array([2.0], F32), "beta": np.array([3.0], F32)}
    kw = dict(g=g, b=b) if model == "lca" else {}
    rec = em.accumulator_records(v, z, t, np.ones((1, 1), F32), [999.0], 20.0, DT, L.BOUNDARY_WEIBULL, bp, **kw)
    mdl = L.MODEL_RACE if model == "race" else L.MODEL_LCA
    job = em.make_job(mdl, L.BOUNDARY_WEIBULL, P, rec, em.uniform_offsets(1, N), delta_t=DT, max_t=20.0, seed=9,
                      smooth_unif=False)
    job.launch()
    torch.cuda.synchronize()
    rg, cg = job.rts[:N].cpu().numpy(), job.choices[:N].cpu().numpy()
    table = E.boundary_table("weibull_cdf", bp, 1, 20.0, DT)
    rc_, cc_, _ = cpu_baseline.em_simulate(1 if model == "race" else 2, v, z, np.full((1, P), E.sqrt_dt32(DT)), t,
                                           [999.0], [20.0], table, np.array([0, N], np.int64), dt=DT, seed=77, **kw)
    _compare_samples(rg, cg, rc_, cc_, range(P))
    assert BF.kernel_boundary_code(BF.weibull_cdf) == L.BOUNDARY_WEIBULL


def test_same_seed_same_output_different_seed_different(torch_gpu):
    import torch

    L, em = _ssm()
    rec = em.ddm_records([0.5], [1.0], [0.5], [0.3], [999.0], [1.0], 20.0, DT, L.BOUNDARY_CONSTANT, {"a": [1.0]})
    outs = []
    for seed in (5, 5, 6):
        job = em.make_job(L.MODEL_DDM, L.BOUNDARY_CONSTANT, 1, rec, em.uniform_offsets(1, 5000), delta_t=DT,
                          max_t=20.0, seed=seed, smooth_unif=True)
        job.launch()
        torch.cuda.synchronize()
        outs.append(job.rts[:5000].cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])


# ----------------------------------------------------------------------------- full DDM and Ornstein
def _run_gpu_diffusion(model, v, a, z, t, bp, bname, N, normals=None, unif=None, smooth=False, seed=1,
                       stride=None, **extra):
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    K = v.shape[0]
    fun = {"constant": BF.constant, "angle": BF.angle}[bname]
    code = BF.kernel_boundary_code(fun)
    rec = em.ddm_records(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), 20.0, DT, code, bp, **extra,
                         stride=stride or L.SET_STRIDE)
    job = em.make_job(model, code, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0, seed=seed,
                      smooth_unif=smooth, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    return job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy(), job.steps()


def _variability(K, seed):
    rng = np.random.default_rng(seed)
    return {"sz": rng.uniform(1e-3, 0.2, K).astype(F32), "sv": rng.uniform(1e-3, 2.0, K).astype(F32),
            "st": rng.uniform(1e-3, 0.25, K).astype(F32)}


@pytest.mark.parametrize("bname", ["constant", "angle"])
def test_full_ddm_replay_matches_oracle(torch_gpu, bname):
    L, _ = _ssm()
    K, N, L_ = 12, 400, 1031
    print(f"[case] full_ddm replay {bname}", flush=True)
    v, a, z, t, bp = _ddm_case(bname, K, 31)
    var = _variability(K, 32)
    rng = np.random.default_rng(4)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=(K * N, 3)).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 boundary=bname, bparams=bp, smooth_unif=True, normals=normals, uniforms=unif,
                                 **var)
    r1, c1, steps = _run_gpu_diffusion(L.MODEL_FULL_DDM, v, a, z, t, bp, bname, N, normals, unif, True, **var)
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert steps == int(s0.sum())


@pytest.mark.parametrize("bname", ["constant", "angle"])
def test_ornstein_replay_matches_oracle(torch_gpu, bname):
    L, _ = _ssm()
    K, N, L_ = 12, 400, 1031
    print(f"[case] ornstein replay {bname}", flush=True)
    v, a, z, t, bp = _ddm_case(bname, K, 41)
    g = np.random.default_rng(42).uniform(-1.0, 1.0, K).astype(F32)
    rng = np.random.default_rng(5)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 boundary=bname, bparams=bp, smooth_unif=True, normals=normals, uniforms=unif, g=g)
    r1, c1, steps = _run_gpu_diffusion(L.MODEL_ORNSTEIN, v, a, z, t, bp, bname, N, normals, unif, True, g=g)
    assert np.array_equal(c0, c1)
    assert np.all(np.abs(r0 - r1) <= DT)
    assert np.mean(r0 == r1) > 0.999
    assert steps == int(s0.sum())


def test_variants_reduce_to_the_ddm_natively(torch_gpu):
    """Same seed, native RNG: the full DDM with zero variability and the Ornstein model with
    zero leak (z = 1/2) walk the DDM's normals and give its output bit for bit."""
    L, _ = _ssm()
    K, N = 8, 3000
    v, a, _, t, bp = _ddm_case("angle", K, 51)
    z = np.full(K, 0.5, F32)
    zeros = np.zeros(K, F32)
    ref = _run_gpu_diffusion(L.MODEL_DDM, v, a, z, t, bp, "angle", N, smooth=True, seed=7)
    full = _run_gpu_diffusion(L.MODEL_FULL_DDM, v, a, z, t, bp, "angle", N, smooth=True, seed=7, sz=zeros, sv=zeros,
                              st=zeros)
    ou = _run_gpu_diffusion(L.MODEL_ORNSTEIN, v, a, z, t, bp, "angle", N, smooth=True, seed=7, g=zeros)
    for got in (full, ou):
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and got[2] == ref[2]


@pytest.mark.parametrize("model", ["full_ddm", "ornstein"])
def test_variants_native_match_numpy_oracle_statistically(torch_gpu, model):
    L, _ = _ssm()
    N = 100_000
    v, a, z, t = (np.array([x], F32) for x in (0.6, 1.4, 0.45, 0.3))
    bp = {"a": a}
    extra = ({"sz": np.array([0.15], F32), "sv": np.array([0.8], F32), "st": np.array([0.2], F32)}
             if model == "full_ddm" else {"g": np.array([0.7], F32)})
    kind = L.MODEL_FULL_DDM if model == "full_ddm" else L.MODEL_ORNSTEIN
    r1, c1, _ = _run_gpu_diffusion(kind, v, a, z, t, bp, "constant", N, smooth=True, seed=11, **extra)
    r0, c0, _ = E.ddm_flexbound(v, a, z, t, [999.0], [1.0], n_samples=N, delta_t=DT, smooth_unif=True,
                                rng=np.random.default_rng(12), **extra)
    p0, p1 = np.mean(c0 == 1), np.mean(c1 == 1)
    assert abs(p0 - p1) <= 3 * math.sqrt(2 * p0 * (1 - p0) / N)
    for ch in (-1, 1):
        x, y = r0[c0 == ch], r1[c1 == ch]
        assert stats.ks_2samp(x, y).pvalue > 0.01
        _quantile_check(x, y)


# ----------------------------------------------------------------------------- Levy flight
@pytest.mark.parametrize("bname", ["constant", "angle"])
def test_levy_replay_matches_oracle(torch_gpu, bname):
    """Replay rows hold alpha-stable variates (drawn per set's alpha by the oracle's sampler)."""
    L, _ = _ssm()
    K, N, L_ = 12, 400, 1031
    print(f"[case] levy replay {bname}", flush=True)
    v, a, z, t, bp = _ddm_case(bname, K, 61)
    rng = np.random.default_rng(62)
    alpha = rng.uniform(1.0, 2.0, K).astype(F32)
    alpha[0] = 1.0
    alpha[1] = 2.0
    noise = E.draw_stable(rng, K * N * L_, np.repeat(alpha, N * L_)).reshape(K * N, L_)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 boundary=bname, bparams=bp, smooth_unif=True, normals=noise, uniforms=unif,
                                 alpha=alpha)
    r1, c1, steps = _run_gpu_diffusion(L.MODEL_LEVY, v, a, z, t, bp, bname, N, noise, unif, True, alpha=alpha)
    assert np.array_equal(c0, c1)
    assert np.all(np.abs(r0 - r1) <= DT)
    assert np.mean(r0 == r1) > 0.999
    assert steps == int(s0.sum())


@pytest.mark.parametrize("alpha", [1.2, 1.7])
def test_levy_native_matches_numpy_oracle_statistically(torch_gpu, alpha):
    L, _ = _ssm()
    N = 100_000
    print(f"[case] levy native alpha={alpha}", flush=True)
    v, a, z, t = (np.array([x], F32) for x in (0.7, 1.3, 0.45, 0.25))
    al = np.array([alpha], F32)
    r1, c1, _ = _run_gpu_diffusion(L.MODEL_LEVY, v, a, z, t, {"a": a}, "constant", N, smooth=True, seed=13,
                                   alpha=al)
    r0, c0, _ = E.ddm_flexbound(v, a, z, t, [999.0], [1.0], n_samples=N, delta_t=DT, smooth_unif=True,
                                rng=np.random.default_rng(14), alpha=al)
    p0, p1 = np.mean(c0 == 1), np.mean(c1 == 1)
    assert abs(p0 - p1) <= 3 * math.sqrt(2 * p0 * (1 - p0) / N)
    for ch in (-1, 1):
        x, y = r0[c0 == ch], r1[c1 == ch]
        assert stats.ks_2samp(x, y).pvalue > 0.01
        _quantile_check(x, y)


def test_levy_native_mean_exit_time_matches_getoor(torch_gpu):
    """E[T] = (1 - x^2)^(alpha/2) / Gamma(1 + alpha) on (-1, 1) for the symmetric stable process
    (Getoor 1961): pins the in-kernel sampler's law and step scale without the oracle."""
    L, _ = _ssm()
    N = 100_000
    alphas = np.array([1.0, 1.4, 1.8], F32)
    K = alphas.size
    ones = np.ones(K, F32)
    r, _, _ = _run_gpu_diffusion(L.MODEL_LEVY, np.zeros(K, F32), ones, np.full(K, 0.5, F32), np.zeros(K, F32),
                                 {"a": ones}, "constant", N, seed=21, alpha=alphas)
    for k, al in enumerate(alphas):
        rk = r[k * N:(k + 1) * N]
        exact = 1.0 / math.gamma(1 + float(al))
        assert abs(rk.mean() - exact) < 0.07 * exact + 4 * rk.std() / math.sqrt(N), (al, rk.mean(), exact)


# ----------------------------------------------------------------------------- racing diffusion
def _rdm_case(K, P, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(0.0, 2.5, (K, P)).astype(F32)
    A = rng.uniform(0.0, 1.0, K).astype(F32)
    b = rng.uniform(0.5, 3.0, K).astype(F32)
    t = rng.uniform(0.0, 2.0, K).astype(F32)
    return v, A, b, t


def _run_gpu_rdm(v, A, b, t, dl, N, normals=None, unif=None, smooth=False, seed=1):
    import torch

    L, em = _ssm()
    K, P = v.shape
    rec = em.accumulator_records(v, A.reshape(K, 1), t, np.ones((K, 1), F32), dl, 20.0, DT, L.BOUNDARY_CONSTANT,
                                 {"a": b})
    job = em.make_job(L.MODEL_RDM, L.BOUNDARY_CONSTANT, P, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=20.0,
                      seed=seed, smooth_unif=smooth, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    return job.rts[: K * N].cpu().numpy(), job.choices[: K * N].cpu().numpy(), job.steps()


@pytest.mark.parametrize("P", [3, 6])
def test_racing_diffusion_replay_matches_oracle(torch_gpu, P):
    K, N, L_ = 10, 300, 3001
    print(f"[case] racing diffusion replay P={P}", flush=True)
    v, A, b, t = _rdm_case(K, P, 70 + P)
    dl = np.full(K, 999.0, F32)
    dl[1] = F32(t[1] + 0.3)  # a deadline that cuts walks off
    dl[2] = F32(0.0)         # deadline <= 0: every trajectory omitted
    rng = np.random.default_rng(72)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=(K * N, P + 1)).astype(F32)
    r0, c0, s0 = E.racing_diffusion(v, A, b, t, np.ones(K, F32), dl, n_samples=N, delta_t=DT, smooth_unif=True,
                                    normals=normals, uniforms=unif)
    r1, c1, steps = _run_gpu_rdm(v, A, b, t, dl, N, normals, unif, True)
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert steps == int(s0.sum())
    assert np.all(c1[2 * N:3 * N] == -1) and np.any(c1[N:2 * N] == -1)


def test_racing_diffusion_native_matches_numpy_oracle_statistically(torch_gpu):
    N = 100_000
    v = np.array([[1.2, 0.9, 0.6]], F32)
    A, b, t = np.array([0.5], F32), np.array([1.4], F32), np.array([0.2], F32)
    r1, c1, _ = _run_gpu_rdm(v, A, b, t, np.array([999.0], F32), N, smooth=True, seed=17)
    r0, c0, _ = E.racing_diffusion(v, A, b, t, [1.0], [999.0], n_samples=N, delta_t=DT, smooth_unif=True,
                                   rng=np.random.default_rng(18))
    _compare_samples(r1, c1, r0, c0, (0, 1, 2))


# ----------------------------------------------------------------------------- two-stage DDMs
def _two_stage_case(K, seed):
    rng = np.random.default_rng(seed)
    v = [rng.uniform(-3, 3, K).astype(F32) for _ in range(3)]
    z = [rng.uniform(0.2, 0.8, K).astype(F32) for _ in range(3)]
    t = rng.uniform(0.0, 1.0, K).astype(F32)
    return v, z, t


def _run_gpu_two_stage(model, v, z, t, dl, bp, bname, N, normals=None, unif=None, smooth=False, seed=1,
                       max_t=20.0):
    import torch

    L, em = _ssm()
    from ssmhip import boundary_functions as BF

    K = t.shape[0]
    fun = {"constant": BF.constant, "angle": BF.angle}[bname]
    code = BF.kernel_boundary_code(fun)
    rec = em.two_stage_records(*v, *z, t, dl, np.ones(K, F32), max_t, DT, code, bp)
    job = em.make_job(model, code, 1, rec, em.uniform_offsets(K, N), delta_t=DT, max_t=max_t, seed=seed,
                      smooth_unif=smooth, replay_normals=normals, replay_uniforms=unif)
    job.launch()
    torch.cuda.synchronize()
    planes = job.rts.cpu().numpy()
    return planes, job.choices[: K * N].cpu().numpy(), job.steps()


@pytest.mark.parametrize("kind,bname", [("par2", "constant"), ("par2", "angle"), ("seq2", "constant"),
                                        ("seq2", "angle")])
def test_two_stage_replay_matches_oracle(torch_gpu, kind, bname):
    L, _ = _ssm()
    K, N, L_ = 12, 300, 4099
    print(f"[case] two-stage replay {kind} {bname}", flush=True)
    v, z, t = _two_stage_case(K, 81)
    _, a, _, _, bp = _ddm_case(bname, K, 82)
    z[1][0], z[2][0] = F32(1.2), F32(-0.2)  # seq2: low starts outside the bounds, prior choice without a walk
    dl = np.full(K, 999.0, F32)
    dl[1] = F32(t[1] + 0.4)
    rng = np.random.default_rng(83)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=(K * N, 4)).astype(F32)
    r0, c0, s0, rh, rl = E.two_stage(kind, *v, *z, t, dl, np.ones(K, F32), n_samples=N, delta_t=DT, boundary=bname,
                                     bparams=bp, smooth_unif=True, normals=normals, uniforms=unif)
    model = L.MODEL_PAR2 if kind == "par2" else L.MODEL_SEQ2
    planes, c1, steps = _run_gpu_two_stage(model, v, z, t, dl, bp, bname, N, normals, unif, True)
    n = K * N
    assert np.array_equal(c0, c1)
    assert np.all(np.abs(r0 - planes[:n]) <= DT) and np.mean(r0 == planes[:n]) > 0.999
    assert steps == int(s0.sum())
    if kind == "par2":
        assert np.array_equal(rh, planes[n:2 * n]) and np.array_equal(rl, planes[2 * n:3 * n])
    assert set(np.unique(c1)) == {0, 1, 2, 3}


def test_seq2_replay_at_max_t_matches_oracle(torch_gpu):
    """Walks that never leave wide bounds reach max_t: the low choice from its prior."""
    L, _ = _ssm()
    K, N, L_ = 2, 500, 2003
    v = [np.zeros(K, F32)] * 3
    z = [np.full(K, 0.5, F32), np.array([0.25, 0.7], F32), np.array([0.8, 0.3], F32)]
    t = np.array([0.1, 0.2], F32)
    bp = {"a": np.full(K, 50.0, F32)}
    rng = np.random.default_rng(84)
    normals = (0.05 * rng.standard_normal((K * N, L_))).astype(F32)
    unif = rng.uniform(size=(K * N, 4)).astype(F32)
    dl = np.full(K, 999.0, F32)
    r0, c0, s0, _, _ = E.two_stage("seq2", *v, *z, t, dl, np.ones(K, F32), n_samples=N, max_t=2.0, delta_t=DT,
                                   bparams=bp, normals=normals, uniforms=unif)
    planes, c1, steps = _run_gpu_two_stage(L.MODEL_SEQ2, v, z, t, dl, bp, "constant", N, normals, unif, max_t=2.0)
    assert np.array_equal(c0, c1) and np.array_equal(r0, planes[:K * N]) and steps == int(s0.sum())
    assert np.all(planes[:K * N] > 2.0)


@pytest.mark.parametrize("kind", ["par2", "seq2"])
def test_two_stage_native_matches_numpy_oracle_statistically(torch_gpu, kind):
    L, _ = _ssm()
    N = 100_000
    v = [np.array([x], F32) for x in (0.9, 0.6, -0.8)]
    z = [np.array([x], F32) for x in (0.55, 0.4, 0.6)]
    t = np.array([0.3], F32)
    bp = {"a": np.array([1.1], F32)}
    model = L.MODEL_PAR2 if kind == "par2" else L.MODEL_SEQ2
    planes, c1, _ = _run_gpu_two_stage(model, v, z, t, np.array([999.0], F32), bp, "constant", N, smooth=True,
                                       seed=19)
    r0, c0, _, _, _ = E.two_stage(kind, *v, *z, t, [999.0], [1.0], n_samples=N, delta_t=DT, bparams=bp,
                                  smooth_unif=True, rng=np.random.default_rng(20))
    _compare_samples(planes[:N], c1, r0, c0, (0, 1, 2, 3))


def test_legacy_ddm_replay_matches_oracle(torch_gpu):
    L, _ = _ssm()
    K, N, L_ = 12, 400, 1031
    print("[case] legacy ddm replay", flush=True)
    v, a, z, t, _ = _ddm_case("constant", K, 91)
    rng = np.random.default_rng(92)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 smooth_unif=True, normals=normals, uniforms=unif, legacy=True)
    r1, c1, steps = _run_gpu_diffusion(L.MODEL_DDM_LEGACY, v, a, z, t, {"a": a}, "constant", N, normals, unif, True)
    assert np.array_equal(c0, c1) and np.array_equal(r0, r1) and steps == int(s0.sum())


@pytest.mark.parametrize("model", ["ddm", "full_ddm", "ornstein", "levy", "legacy"])
def test_compact_records_give_the_same_walks(torch_gpu, model):
    """16-float records (set_stride 16) of the single-particle diffusions: same results, bit
    for bit, as the 64-float records; the other models refuse them."""
    import ctypes as C

    L, em = _ssm()
    from ssmhip import _device as D

    print(f"[case] compact records {model}", flush=True)
    K, N = 24, 500
    v, a, z, t, bp = _ddm_case("angle", K, 41)
    mdl = {"ddm": L.MODEL_DDM, "full_ddm": L.MODEL_FULL_DDM, "ornstein": L.MODEL_ORNSTEIN, "levy": L.MODEL_LEVY,
           "legacy": L.MODEL_DDM_LEGACY}[model]
    extra = {"full_ddm": _variability(K, 42), "ornstein": {"g": np.full(K, 0.3, F32)},
             "levy": {"alpha": np.full(K, 1.6, F32)}}.get(model, {})
    if model == "legacy":
        z = np.full(K, 0.5, F32)
        bp = {"a": a}
    bname = "constant" if model == "legacy" else "angle"
    full = _run_gpu_diffusion(mdl, v, a, z, t, bp, bname, N, smooth=True, seed=5, **extra)
    compact = _run_gpu_diffusion(mdl, v, a, z, t, bp, bname, N, smooth=True, seed=5, stride=L.SET_STRIDE_SINGLE,
                                 **extra)
    assert np.array_equal(full[0], compact[0]) and np.array_equal(full[1], compact[1]) and full[2] == compact[2]
    if model == "ddm":
        # a race with 16-float records is refused before anything is launched
        rec = em.accumulator_records(np.ones((2, 2), F32), np.full((2, 2), 0.3, F32), np.zeros(2, F32),
                                     np.ones((2, 2), F32), np.full(2, 999, F32), 20.0, DT, L.BOUNDARY_CONSTANT,
                                     {"a": np.ones(2, F32)})
        job = em.make_job(L.MODEL_RACE, L.BOUNDARY_CONSTANT, 2, rec, em.uniform_offsets(2, 10), delta_t=DT,
                          max_t=20.0, seed=1, smooth_unif=False)
        job.prob.set_stride = L.SET_STRIDE_SINGLE
        with pytest.raises(ValueError, match="set_stride"):
            job.launch()


# ----------------------------------------------------------------------------- metadata['trajectory']
def test_trajectory_metadata_matches_oracle_in_replay_mode(torch_gpu):
    """The drop-ins' metadata['trajectory'] (trial 0, sample 0, every step) equals the oracle's
    recording of the same walk, for the DDM with an angle bound, the full DDM, race and LCA."""
    from ssmhip import boundary_functions as BF
    from ssmhip import cssm

    K, N, L_ = 3, 20, 257
    rng = np.random.default_rng(71)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    unif = rng.uniform(size=(K * N, 3)).astype(F32)
    v = np.array([0.6, -0.4, 1.0], F32)
    a = np.array([1.2, 1.0, 1.5], F32)
    z = np.array([0.45, 0.55, 0.5], F32)
    t = np.array([0.2, 0.3, 0.1], F32)
    dl, s1 = np.full(K, 999, F32), np.ones(K, F32)
    th = np.array([0.2, 0.5, 0.1], F32)
    print("[case] trajectory metadata, replay", flush=True)
    out = cssm.ddm_flexbound(v, a, z, t, dl, s1, n_samples=N, n_trials=K, boundary_fun=BF.angle,
                             boundary_params={"a": a, "theta": th}, replay_normals=normals,
                             replay_uniforms=unif[:, 0].copy())
    *_, tr = E.ddm_flexbound(v, a, z, t, dl, s1, n_samples=N, boundary="angle", bparams={"a": a, "theta": th},
                             normals=normals, uniforms=unif[:, 0].copy(), record0=True)
    assert np.array_equal(out["metadata"]["trajectory"], tr)
    var = {"sz": np.full(K, 0.1, F32), "sv": np.full(K, 0.5, F32), "st": np.full(K, 0.05, F32)}
    out = cssm.full_ddm(v, a, z, t, var["sz"], var["sv"], var["st"], dl, s1, n_samples=N, n_trials=K,
                        replay_normals=normals, replay_uniforms=unif)
    *_, tr = E.ddm_flexbound(v, a, z, t, dl, s1, n_samples=N, normals=normals, uniforms=unif, record0=True, **var)
    assert np.array_equal(out["metadata"]["trajectory"], tr)
    P = 3
    vr = rng.uniform(0.5, 2.0, (K, P)).astype(F32)
    zr = rng.uniform(0.0, 0.4, (K, P)).astype(F32)
    ar = np.array([1.5, 2.0, 1.2], F32)
    for model in ("race", "lca"):
        kw = dict(g=np.full((K, 1), 0.3, F32), b=np.full((K, 1), 0.2, F32)) if model == "lca" else {}
        fn = cssm.race_model if model == "race" else cssm.lca
        out = fn(v=vr, z=zr, t=t[:, None], s=np.ones((K, 1), F32), deadline=dl, n_samples=N, n_trials=K,
                 boundary_fun=BF.constant, boundary_params={"a": ar}, replay_normals=normals, **kw)
        *_, tr = E.accumulators(model, vr, zr, t, np.ones((K, 1), F32), dl, n_samples=N, bparams={"a": ar},
                                normals=normals, record0=True, **kw)
        assert np.array_equal(out["metadata"]["trajectory"], tr), model


def test_trajectory_metadata_in_native_mode_is_walk_0(torch_gpu):
    """Native draws: the trajectory ends where trial 0, sample 0 ended (its sign is the choice,
    its length the rt's step count) and stays inside the bounds until then."""
    from ssmhip import cssm

    out = cssm.ddm_flexbound(np.array([0.8], F32), np.array([1.0], F32), np.array([0.5], F32), np.array([0.3], F32),
                             np.array([999.0], F32), np.array([1.0], F32), n_samples=1000, n_trials=1,
                             random_state=3)
    tr = out["metadata"]["trajectory"][:, 0]
    n = int(np.sum(tr != -999.0)) - 1
    assert n >= 1 and np.all(tr[n + 1:] == -999.0)
    assert np.all(np.abs(tr[:n]) <= 1.0) and abs(tr[n]) > 1.0
    assert np.sign(tr[n]) == out["choices"][0, 0, 0]
    t_walk = F32(0)
    for _ in range(n):
        t_walk = F32(t_walk + F32(DT))
    assert out["rts"][0, 0, 0] == F32(t_walk + F32(0.3))  # smooth_unif off in the drop-in's default
