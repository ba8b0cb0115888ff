"""Pins of the EM oracle (CPU): closed forms, identities, and the C++ restatement.

The NumPy transcription (oracle/em_numpy.py) is checked against mathematics
that fixes the answer independently of its code: the Wiener absorption
probability and mean exit time (Siegmund-corrected for the Euler overshoot),
the noise-free crossing step, one-accumulator first passage, the
LCA(g=b=0) == race identity, and the LBA closed form. The C++/OpenMP
restatement (oracle/cpu_baseline.cpp) is pinned bit-exactly to it in replay.
"""

import math

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import cpu_baseline, ddm_analytic
from oracle import em_numpy as E

F32 = np.float32
DT = 0.001
SQDT = math.sqrt(DT)


def test_time_grid_identity():
    # t_s[ix] == float32(ix) * float32(dt) exactly: lets kernels rebuild t_s in registers
    for max_t, dt in [(20.0, 0.001), (5.0, 0.002), (3.0, 0.0005)]:
        ts = E.time_grid(max_t, dt)
        ix = np.arange(ts.shape[0], dtype=np.float32)
        assert np.array_equal(ts, (ix * F32(dt)).astype(F32))
        assert ts.shape[0] == E.n_steps(max_t, dt)


def test_analytic_density_integrates_to_absorption_probability():
    v, a, z = 1.0, 1.5, 0.5
    up, _ = integrate.quad(lambda t: ddm_analytic.density(np.array([t]), v, a, z, 1)[0], 0, 30, limit=200)
    lo, _ = integrate.quad(lambda t: ddm_analytic.density(np.array([t]), v, a, z, -1)[0], 0, 30, limit=200)
    assert abs(up - ddm_analytic.p_upper(v, a, z)) < 1e-6
    assert abs(up + lo - 1.0) < 1e-6
    # closed forms at v=1, a=1.5, z=0.5, t=0.2
    assert abs(ddm_analytic.p_upper(1.0, 1.5, 0.5) - 0.952574) < 1e-6
    assert abs(ddm_analytic.mean_decision_time(1.0, 1.5, 0.5) - 1.35772) < 1e-5


@pytest.mark.parametrize("v,a,z", [(1.0, 1.5, 0.5), (-0.7, 1.0, 0.35)])
def test_ddm_oracle_matches_wiener_closed_form(v, a, z):
    n = 20000
    rng = np.random.default_rng(11)
    rts, ch, _ = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [1.0], n_samples=n, delta_t=DT, rng=rng)
    # Euler overshoot: both bounds move out by beta*sqrt(dt); start keeps its absolute position
    a_eff = a + ddm_analytic.SIEGMUND_BETA * SQDT
    y0 = (2 * z - 1) * a
    z_eff = (y0 + a_eff) / (2 * a_eff)
    p = ddm_analytic.p_upper(v, a_eff, z_eff)
    se = math.sqrt(p * (1 - p) / n)
    assert abs(np.mean(ch == 1) - p) < 4 * se
    mt = ddm_analytic.mean_decision_time(v, a_eff, z_eff)
    se_t = np.std(rts) / math.sqrt(n)
    assert abs(np.mean(rts) - mt) < 4 * se_t + 0.003


def test_ddm_oracle_noise_free_crossing_step():
    # s = 0: y_k = y0 + k*v*dt (float32 sums); the walk stops at the first k with y_k > a
    v, a, z = 0.8, 1.0, 0.5
    rts, ch, steps = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [0.0], n_samples=3, delta_t=DT,
                                     rng=np.random.default_rng(0))
    y = F32(0.0)
    k = 0
    while -F32(a) <= y <= F32(a):
        y = F32(y + F32(F32(v) * F32(DT)))
        k += 1
    assert np.all(steps == k)
    assert abs(k - a / (v * DT)) <= 2
    assert np.all(ch == 1)


def test_race_single_accumulator_is_inverse_gaussian():
    # one accumulator far from the clamp at 0: first passage of a drifted Wiener
    # process; mean = (a_eff - x0)/v, var = (a_eff - x0)/v^3
    v, a, z = 2.0, 2.0, 0.9
    n = 20000
    rts, ch, _ = E.accumulators("race", np.full((1, 1), v, F32), [[z]], [0.0], [1.0], [999.0], n_samples=n,
                                delta_t=DT, bparams={"a": [a]}, rng=np.random.default_rng(5))
    dist = a + ddm_analytic.SIEGMUND_BETA * SQDT - z * a
    assert np.all(ch == 0)
    assert abs(rts.mean() - dist / v) < 4 * rts.std() / math.sqrt(n) + 0.002
    assert abs(rts.var() - dist / v**3) < 0.1 * dist / v**3


def test_lca_without_leak_or_inhibition_equals_race():
    K, P, N, L = 3, 4, 300, 257
    rng = np.random.default_rng(2)
    v = rng.uniform(0, 2.5, (K, P)).astype(F32)
    z = rng.uniform(0, 0.5, (K, P)).astype(F32)
    a = rng.uniform(1, 3, K).astype(F32)
    normals = rng.standard_normal((K * N, L)).astype(F32)
    common = dict(n_samples=N, delta_t=DT, bparams={"a": a}, normals=normals)
    r1, c1, s1 = E.accumulators("race", v, z, np.zeros(K, F32), np.ones((K, 1), F32), np.full(K, 999, F32), **common)
    r2, c2, s2 = E.accumulators("lca", v, z, np.zeros(K, F32), np.ones((K, 1), F32), np.full(K, 999, F32),
                                g=np.zeros(K, F32), b=np.zeros(K, F32), **common)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    assert np.array_equal(r1, r2)


def test_lba_closed_form_choice_probability():
    # start-point range 0: x_j = a/|v_j + sd N_j|; with v >> sd, P(choice 0) = Phi((v0 - v1)/(sd*sqrt 2))
    n = 200000
    rng = np.random.default_rng(4)
    v = np.array([[1.0, 0.9]], F32)
    rts, ch = E.lba(v, [1.0], [0.0], [0.0], [999.0], [0.1], n_samples=n,
                    uniforms=rng.uniform(size=(n, 2)).astype(F32), normals=rng.standard_normal((n, 2)).astype(F32))
    p0 = stats.norm.cdf(0.1 / (0.1 * math.sqrt(2)))
    assert abs(np.mean(ch == 0) - p0) < 4 * math.sqrt(p0 * (1 - p0) / n)
    assert np.all(rts > 0)


def _ddm_cpu_args(v, a, z, t, s, bname, bparams, K):
    table = E.boundary_table(bname, bparams, K, 20.0, DT)
    dl = np.minimum(F32(20.0), (F32(999.0) - t).astype(F32)).astype(F32)
    sst = (E.sqrt_dt32(DT) * s).astype(F32)
    return table, dl, sst


@pytest.mark.parametrize("bname", ["constant", "angle", "weibull_cdf"])
def test_cpu_restatement_replay_matches_numpy_oracle_ddm(bname):
    K, N, L = 6, 200, 311
    rng = np.random.default_rng(9)
    v = rng.uniform(-3, 3, K).astype(F32)
    a = rng.uniform(0.3, 2.5, K).astype(F32)
    z = rng.uniform(0.1, 0.9, K).astype(F32)
    t = rng.uniform(0, 2, K).astype(F32)
    s = np.ones(K, F32)
    bparams = {"a": a, "theta": rng.uniform(-0.1, 1.3, K).astype(F32), "alpha": rng.uniform(0.3, 5, K).astype(F32),
               "beta": rng.uniform(0.3, 7, K).astype(F32)}
    bp = {k: bparams[k] for k in E.BOUNDARIES[bname][1]}
    normals = rng.standard_normal((K * N, L)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), s, n_samples=N, delta_t=DT, boundary=bname,
                                 bparams=bp, smooth_unif=True, normals=normals, uniforms=unif)
    table, dl, sst = _ddm_cpu_args(v, a, z, t, s, bname, bp, K)
    offs = np.arange(K + 1, dtype=np.int64) * N
    r1, c1, steps = cpu_baseline.em_simulate(0, v, z, sst, t, np.full(K, 999, F32), dl, table, offs, dt=DT,
                                             smooth_unif=True, replay_normals=normals, replay_uniforms=unif)
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert steps == int(s0.sum())


@pytest.mark.parametrize("model", ["race", "lca"])
def test_cpu_restatement_replay_matches_numpy_oracle_accumulators(model):
    K, P, N, L = 4, 4, 150, 509
    rng = np.random.default_rng(10)
    v = rng.uniform(0, 2.5, (K, P)).astype(F32)
    z = rng.uniform(0, 0.5, (K, P)).astype(F32)
    a = rng.uniform(1, 3, K).astype(F32)
    t = rng.uniform(0, 2, K).astype(F32)
    g = rng.uniform(-1, 1, K).astype(F32)
    b = rng.uniform(-1, 1, K).astype(F32)
    normals = rng.standard_normal((K * N, L)).astype(F32)
    unif = rng.uniform(size=K * N).astype(F32)
    kw = dict(g=g, b=b) if model == "lca" else {}
    bp = {"a": a, "alpha": rng.uniform(0.3, 5, K).astype(F32), "beta": rng.uniform(0.3, 7, K).astype(F32)}
    r0, c0, s0 = E.accumulators(model, v, z, t, np.ones((K, 1), F32), np.full(K, 999, F32), n_samples=N,
                                delta_t=DT, boundary="weibull_cdf", bparams=bp, smooth_unif=True, normals=normals,
                                uniforms=unif, **kw)
    table = E.boundary_table("weibull_cdf", bp, K, 20.0, DT)
    dl = np.minimum(F32(20.0), (F32(999.0) - t).astype(F32)).astype(F32)
    sst = np.full((K, P), E.sqrt_dt32(DT), F32)
    offs = np.arange(K + 1, dtype=np.int64) * N
    r1, c1, steps = cpu_baseline.em_simulate(1 if model == "race" else 2, v, z, sst, t, np.full(K, 999, F32), dl,
                                             table, offs, dt=DT, smooth_unif=True, replay_normals=normals,
                                             replay_uniforms=unif, **kw)
    assert np.array_equal(c0, c1)
    assert np.array_equal(r0, r1)
    assert steps == int(s0.sum())


def test_cpu_baseline_native_matches_closed_form():
    # MT19937 + Ziggurat path of the C++ restatement against the Wiener closed form (v=1, a=1.5, z=0.5)
    v, a, z = 1.0, 1.5, 0.5
    K, N = 1, 50000
    table = E.boundary_table("constant", {"a": [a]}, K, 20.0, DT)
    offs = np.array([0, N], np.int64)
    rts, ch, _ = cpu_baseline.em_simulate(0, [v], [z], [E.sqrt_dt32(DT)], [0.0], [999.0], [20.0], table, offs,
                                          dt=DT, seed=0)
    a_eff = a + ddm_analytic.SIEGMUND_BETA * SQDT
    p = ddm_analytic.p_upper(v, a_eff, 0.5)
    assert abs(np.mean(ch == 1) - p) < 4 * math.sqrt(p * (1 - p) / N)
    assert abs(rts.mean() - ddm_analytic.mean_decision_time(v, a_eff, 0.5)) < 4 * rts.std() / math.sqrt(N) + 0.003


# ----------------------------------------------------------------------------- full DDM and Ornstein
def test_full_ddm_without_variability_is_the_ddm():
    rng = np.random.default_rng(1)
    K, N, L = 3, 100, 20002  # rows long enough that no walk wraps around its row
    v = rng.uniform(-1, 1, K).astype(F32)
    a = rng.uniform(0.8, 1.5, K).astype(F32)
    z = rng.uniform(0.3, 0.7, K).astype(F32)
    t = rng.uniform(0.1, 0.5, K).astype(F32)
    nz = rng.standard_normal((K * N, L)).astype(F32)
    u3 = rng.uniform(0, 1, (K * N, 3)).astype(F32)
    kw = dict(n_samples=N, delta_t=DT, smooth_unif=True)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), normals=nz[:, 1:],
                                 uniforms=u3[:, 2], **kw)
    zeros = np.zeros(K, F32)
    r1, c1, s1 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), normals=nz, uniforms=u3,
                                 sz=zeros, sv=zeros, st=zeros, **kw)
    # the walk normals of the full model sit one column later (column 0 is the drift draw)
    assert np.array_equal(c0, c1) and np.array_equal(s0, s1) and np.array_equal(r0, r1)


def test_full_ddm_noise_free_variability_closed_forms():
    # s = 0: each walk is deterministic given its drift draw gv and uniforms (u_t, u_z)
    K, N = 1, 64
    v, a, z, t, sz, sv, st = 0.5, 1.0, 0.5, 0.3, 0.2, 0.4, 0.1
    rng = np.random.default_rng(2)
    nz = np.zeros((N, 8), F32)
    nz[:, 0] = rng.standard_normal(N).astype(F32)
    u3 = rng.uniform(0, 1, (N, 3)).astype(F32)
    rts, ch, steps = E.ddm_flexbound([v], [a], [z], [t], [999.0], [0.0], n_samples=N, delta_t=DT, normals=nz,
                                     uniforms=u3, sz=[sz], sv=[sv], st=[st])
    vk = v + sv * nz[:, 0].astype(np.float64)
    y0 = 2 * (u3[:, 1].astype(np.float64) - 0.5) * sz          # start shift around 0
    ttmp = t + 2 * (u3[:, 0].astype(np.float64) - 0.5) * st
    target = np.where(vk > 0, a - y0, -a - y0)                   # distance to the boundary it reaches
    k_exact = np.floor(target / (vk * DT)) + 1                  # first step beyond the boundary
    # exact-arithmetic crossing step; float32 sums over thousands of steps may drift by ~1e-4
    assert np.all(np.abs(steps - k_exact) <= np.maximum(1, 5e-4 * k_exact))
    assert np.all(ch == np.sign(vk))
    # rt = t_particle + t_tmp, the clock being a float32 sum of dt
    clock = np.zeros(N, F32)
    for i in range(int(steps.max())):
        clock = np.where(i < steps, (clock + F32(DT)).astype(F32), clock)
    assert np.allclose(rts, clock.astype(np.float64) + ttmp, atol=1e-6)


def test_ornstein_without_leak_is_the_ddm_at_z_half():
    rng = np.random.default_rng(3)
    K, N, L = 3, 200, 257
    v = rng.uniform(-1, 1, K).astype(F32)
    a = rng.uniform(0.8, 1.5, K).astype(F32)
    z = np.full(K, 0.5, F32)  # start 0 exactly in both start formulas
    t = rng.uniform(0.1, 0.5, K).astype(F32)
    nz = rng.standard_normal((K * N, L)).astype(F32)
    kw = dict(n_samples=N, delta_t=DT, normals=nz)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), **kw)
    r1, c1, s1 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), g=np.zeros(K, F32), **kw)
    assert np.array_equal(c0, c1) and np.array_equal(s0, s1) and np.array_equal(r0, r1)


@pytest.mark.parametrize("g", [0.8, -0.5])
def test_ornstein_noise_free_crossing_matches_the_linear_recurrence(g):
    # s = 0: y_{n+1} = y_n + (v - g y_n) dt, so y_n = v/g + (y0 - v/g)(1 - g dt)^n; the walk stops
    # at the first n with y_n > a. A wrong sign or order of the leak moves that step far.
    v, a, z = 1.0, 0.6, 0.5
    rts, ch, steps = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [0.0], n_samples=2, delta_t=DT,
                                     g=[g], rng=np.random.default_rng(0))
    n = np.arange(1, 20001)
    yn = v / g + (0.0 - v / g) * (1 - g * DT) ** n
    k = int(n[np.argmax(yn > a)])
    assert np.all(np.abs(steps - k) <= 1) and np.all(ch == 1)
    assert abs(k - 0.6 / DT) > 20  # the leak visibly changes the crossing step
