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


def test_ddm_start_point_follows_the_reference_c_precision():
    # ddm_models.pyx:709 `(-1) * b0 + (z * 2 * b0)`: float operands and an int 2, so each C
    # operation rounds to float: fl(-b0 + fl(fl(2 z) b0)). full_ddm (:2305) writes 2.0, a
    # double: the exact -b0 + 2 z b0 rounded once. Both against exact rationals.
    from fractions import Fraction as Q

    from em_cases import round32

    rng = np.random.default_rng(7)
    z = rng.uniform(0.1, 0.9, 300).astype(F32)
    b0 = rng.uniform(0.3, 3.0, 300).astype(F32)
    z[:4] = [0.5, 0.1, 0.9, 0.3]
    b0[:4] = [1.5, 1.5, 0.7, 2.9]
    seq = E.ddm_start32(z, b0)
    dbl = E.ddm_start32(z, b0, double_literal=True)
    for i in range(z.shape[0]):
        zq, bq = Q(float(z[i])), Q(float(b0[i]))
        prod = round32(Q(float(round32(2 * zq))) * bq)
        assert seq[i] == round32(-bq + Q(float(prod))), i
        assert dbl[i] == round32(-bq + 2 * zq * bq), i
    # the forms differ on about half of such pairs, so a mix-up fails the loop above
    assert 0.3 < np.mean(seq != dbl) < 0.8
    assert seq[0] == dbl[0] == 0.0


def test_ddm_oracle_walk_starts_from_the_float_start_point():
    # v = 0, s = 0: y stays at its start, so the walk ends at the first boundary index with
    # b[ix] < |y0|. On sets where that index depends on the start's rounding, the plain DDM
    # oracle must end where the float start fl(-b0 + fl(2 z b0)) (exact rationals) puts it.
    from fractions import Fraction as Q

    from em_cases import round32, start_sensitive_ddm_cases

    v, a, z, th, _ = start_sensitive_ddm_cases(2000)
    assert v.shape[0] >= 50
    K = v.shape[0]
    _, ch, steps = E.ddm_flexbound(v, a, z, np.zeros(K, F32), np.full(K, 999, F32), np.zeros(K, F32), n_samples=1,
                                   delta_t=DT, boundary="angle", bparams={"a": a, "theta": th},
                                   normals=np.zeros((K, 1), F32))
    table = E.boundary_table("angle", {"a": a, "theta": th}, K, 20.0, DT)
    for i in range(K):
        zq, bq = Q(float(z[i])), Q(float(a[i]))
        y0 = round32(-bq + Q(float(round32(Q(float(round32(2 * zq))) * bq))))
        assert steps[i] == int(np.argmax(table[i] < abs(y0))), i
        assert ch[i] == np.sign(y0), i


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


@pytest.mark.parametrize("angle", [False, True])
def test_lba_float32_transcription_stays_within_ulps_of_the_float64_reference_form(angle):
    # the kernel computes the LBA in float32 (lba()); the reference in NumPy float64 (lba64()).
    # On the same draws: identical choices and rts within the float32 error bound
    # (em_cases.lba_float32_deviation_ok; vanilla: at most 5 ulps measured; angle: more where
    # vs + tan(theta) cancels and magnifies the float32 rounding of tan(theta)).
    from em_cases import lba_float32_deviation_ok

    K, N, P = 32, 20000, 3
    rng = np.random.default_rng(12)
    v = rng.uniform(0.2, 2.5, (K, P)).astype(F32)
    a = rng.uniform(0.5, 2.0, K).astype(F32)
    z = (a * rng.uniform(0.05, 0.95, K)).astype(F32)
    t = rng.uniform(0.0, 1.0, K).astype(F32)
    sd = np.full(K, 0.1, F32)
    th = rng.uniform(-0.1, 0.6, K).astype(F32) if angle else None
    u = rng.uniform(size=(K * N, P)).astype(F32)
    g = rng.standard_normal((K * N, P)).astype(F32)
    kw = dict(n_samples=N, uniforms=u, normals=g)
    tan32 = None if th is None else np.tan(th).astype(F32)
    tan64 = None if th is None else np.tan(th.astype(np.float64))
    r32, c32 = E.lba(v, a, z, t, np.full(K, 999, F32), sd, tan_theta=tan32, **kw)
    r64, c64 = E.lba64(v, a, z, t, np.full(K, 999, F32), sd, tan_theta=tan64, **kw)
    share, within = lba_float32_deviation_ok(r32, c32, r64, c64, v, sd, t, u, g, N, tan32, tan64, a=a, z=z)
    assert share < 1e-5 and within
    if not angle:
        assert np.max(np.abs(r32.view(np.int32).astype(np.int64) - r64.view(np.int32))) <= 8


def test_lba64_choice_probability_closed_form():
    # start-point range 0 and v >> sd: choice 0 iff v0 + sd g0 > v1 + sd g1, P = Phi((v0 - v1)/(sd sqrt 2))
    n = 200000
    rng = np.random.default_rng(14)
    v = np.array([[1.0, 0.9]], F32)
    rts, ch = E.lba64(v, [1.0], [0.0], [0.0], [999.0], [0.1], n_samples=n,
                      uniforms=rng.uniform(size=(n, 2)).astype(F32), normals=rng.standard_normal((n, 2)).astype(F32))
    p0 = stats.norm.cdf(0.1 / (0.1 * math.sqrt(2)))
    assert abs(np.mean(ch == 0) - p0) < 4 * math.sqrt(p0 * (1 - p0) / n)


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


# ----------------------------------------------------------------------------- Levy flight
@pytest.mark.parametrize("alpha", [1.0, 1.25, 1.6, 2.0])
def test_stable_sampler_matches_the_stable_law(alpha):
    """CMS draws against SciPy's symmetric stable law (Cauchy at 1, N(0, 2) at 2); a wrong
    exponent or a swapped cosine moves the law far beyond the KS bound."""
    x = E.draw_stable(np.random.default_rng(int(alpha * 100)), 3000, F32(alpha)).astype(np.float64)
    if alpha == 2.0:
        p = stats.kstest(x / math.sqrt(2.0), "norm").pvalue
    elif alpha == 1.0:
        p = stats.kstest(x, "cauchy").pvalue
    else:
        p = stats.kstest(x, stats.levy_stable(alpha, 0.0).cdf).pvalue
    assert p > 1e-3, p


def test_stable_sampler_special_values():
    u = np.array([0.0, 0.3, -1.2], F32)
    w = np.array([1.0, 0.5, 2.0], F32)
    # alpha = 1: tan(u) whatever w; alpha = 2: 2 sin(u) sqrt(w) (sin 2u / sqrt(cos u) * sqrt(w / cos u))
    assert np.allclose(E.stable_cms(u, w, F32(1.0)), np.tan(u), rtol=1e-6)
    assert np.allclose(E.stable_cms(u, w, F32(2.0)), 2 * np.sin(u) * np.sqrt(w), rtol=1e-5, atol=1e-7)


def test_levy_scale_is_the_self_similar_step():
    # dt^(1/alpha) s: the square-root scale at alpha = 2, dt itself at alpha = 1
    assert E.levy_scale32([1.0], [2.0], DT)[0] == E.sqrt_dt32(DT)
    assert E.levy_scale32([2.0], [1.0], DT)[0] == F32(2 * np.float64(F32(DT)))


def test_levy_at_alpha_two_walks_as_the_ddm_on_the_same_draws():
    """At alpha = 2, s = 1 the Levy step scale is float32(sqrt(dt)), and at z = 1/2 both start
    formulas give 0: fed the same increments, the two walks agree exactly."""
    rng = np.random.default_rng(4)
    K, N, L = 3, 200, 257
    v = rng.uniform(-1, 1, K).astype(F32)
    a = rng.uniform(0.8, 1.5, K).astype(F32)
    z = np.full(K, 0.5, F32)
    t = rng.uniform(0.1, 0.5, K).astype(F32)
    nz = rng.standard_normal((K * N, L)).astype(F32)
    kw = dict(n_samples=N, delta_t=DT, normals=nz)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), **kw)
    r1, c1, s1 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), alpha=np.full(K, 2.0, F32), **kw)
    assert np.array_equal(c0, c1) and np.array_equal(s0, s1) and np.array_equal(r0, r1)


def test_levy_native_at_alpha_two_is_the_ddm_with_root_two_noise():
    # S_2 = N(0, 2): in law the alpha = 2 flight is the DDM with s = sqrt(2)
    v, a, z, n = 0.5, 1.2, 0.5, 6000
    r0, c0, _ = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [1.0], n_samples=n, delta_t=DT, alpha=[2.0],
                                rng=np.random.default_rng(7))
    r1, c1, _ = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [math.sqrt(2.0)], n_samples=n, delta_t=DT,
                                rng=np.random.default_rng(8))
    p0, p1 = np.mean(c0 == 1), np.mean(c1 == 1)
    assert abs(p0 - p1) < 4 * math.sqrt(p0 * (1 - p0) * 2 / n)
    assert stats.ks_2samp(r0, r1).pvalue > 1e-3


@pytest.mark.parametrize("alpha,x0", [(1.0, 0.0), (1.3, 0.0), (1.7, 0.4), (2.0, 0.0)])
def test_levy_mean_exit_time_matches_getoor(alpha, x0):
    """Getoor (1961): the symmetric alpha-stable process with exponent |theta|^alpha leaves
    (-1, 1) from x after E[T] = (1 - x^2)^(alpha/2) / Gamma(1 + alpha) on average. This pins the
    sampler's normalisation and the dt^(1/alpha) step scale together (a wrong exponent in
    either shifts E[T] by far more than the bound); the walk is monitored at the steps only,
    which lengthens it by a few per cent."""
    n = 4000
    z = (x0 + 1.0) / 2.0
    r, _, _ = E.ddm_flexbound([0.0], [1.0], [z], [0.0], [999.0], [1.0], n_samples=n, delta_t=DT, alpha=[alpha],
                              rng=np.random.default_rng(int(alpha * 10)))
    exact = (1 - x0 * x0) ** (alpha / 2) / math.gamma(1 + alpha)
    se = r.std() / math.sqrt(n)
    assert abs(r.mean() - exact) < 0.07 * exact + 4 * se, (r.mean(), exact)


# ----------------------------------------------------------------------------- racing diffusion
def _rdm_loop(v, A, b, t, s, dl, u, normals, smooth):
    """One trajectory, scalar float32, in the control flow of race_models.pyx:1200-1255."""
    P = len(v)
    dt = F32(DT)
    sst = F32(E.sqrt_dt32(DT) * F32(s))
    x = [F32(F32(u[j]) * F32(A)) for j in range(P)]
    tp, m, winner, found = F32(0.0), 0, -1, False
    dl_tmp = min(F32(20.0), F32(F32(dl) - F32(t)))
    while not found and tp <= dl_tmp:
        for j in range(P):
            x[j] = F32(x[j] + F32(F32(F32(v[j]) * dt) + F32(sst * normals[m % normals.size])))
            m += 1
            if x[j] >= F32(b):
                found, winner = True, j
                break
        if found:
            break
        tp = F32(tp + dt)
    su = E.smooth_unif32(np.array([tp]), np.array([dl_tmp]), dt, np.array([u[P]]))[0] if smooth else F32(0)
    rt = F32(F32(tp + F32(t)) + su)
    if rt >= F32(dl) or F32(dl) <= 0:
        rt = F32(-999.0)
    if rt == F32(-999.0) or not found:
        return F32(-999.0), -1
    return rt, winner


def test_racing_diffusion_matches_the_scalar_loop():
    rng = np.random.default_rng(12)
    K, N, P, L = 3, 40, 3, 3001
    v = rng.uniform(0.2, 2.5, (K, P)).astype(F32)
    A = rng.uniform(0.0, 1.0, K).astype(F32)
    b = rng.uniform(0.5, 1.5, K).astype(F32)
    t = rng.uniform(0.0, 0.5, K).astype(F32)
    t[2] = F32(0.1)
    dl = np.array([999.0, 999.0, 0.4], F32)  # one set with a deadline that cuts walks off
    normals = rng.standard_normal((K * N, L)).astype(F32)
    unif = rng.uniform(size=(K * N, P + 1)).astype(F32)
    r, c, _ = E.racing_diffusion(v, A, b, t, np.ones(K, F32), dl, n_samples=N, delta_t=DT, smooth_unif=True,
                                 normals=normals, uniforms=unif)
    for i in range(K * N):
        k = i // N
        r1, c1 = _rdm_loop(v[k], A[k], b[k], t[k], 1.0, dl[k], unif[i], normals[i], True)
        assert r[i] == r1 and c[i] == c1, i
    assert np.any(c == -1) and np.any(c >= 0)


def test_racing_diffusion_noise_free_crossing_step():
    # s = 0: accumulator j reaches b after ceil((b - u_j A) / (v_j dt)) steps; the earliest wins and
    # rt counts the steps completed before the winning one
    v = np.array([[1.0, 2.0, 1.5]], F32)
    u = np.array([[0.5, 0.1, 0.9, 0.3]], F32)
    r, c, st = E.racing_diffusion(v, [0.4], [1.0], [0.2], [0.0], [999.0], n_samples=1,
                                  normals=np.zeros((1, 8), F32), uniforms=u)
    need = np.ceil((1.0 - u[0, :3] * 0.4) / (v[0] * DT))
    j = int(np.argmin(need))
    assert c[0] == j == 2 and abs(st[0] - need[j]) <= 1
    assert abs(r[0] - ((st[0] - 1) * DT + 0.2)) < 2e-5


def test_racing_diffusion_one_accumulator_is_inverse_gaussian():
    # A = 0, one accumulator: first passage of a drifted Wiener process through b, minus the one
    # step that is not counted; mean (b + beta sqrt(dt)) / v - dt (Siegmund), variance b / v^3
    v, b, n = 1.5, 1.2, 20000
    r, c, _ = E.racing_diffusion(np.full((1, 1), v, F32), [0.0], [b], [0.0], [1.0], [999.0], n_samples=n,
                                 delta_t=DT, rng=np.random.default_rng(13))
    assert np.all(c == 0)
    mean = (b + ddm_analytic.SIEGMUND_BETA * SQDT) / v - DT
    assert abs(r.mean() - mean) < 4 * math.sqrt(b / v**3 / n)
    assert abs(r.var() - b / v**3) < 0.1 * b / v**3


# ----------------------------------------------------------------------------- two-stage DDMs
def _siegmund(v, a, z):
    a_eff = a + ddm_analytic.SIEGMUND_BETA * SQDT
    z_eff = ((2 * z - 1) * a + a_eff) / (2 * a_eff)
    return ddm_analytic.p_upper(v, a_eff, z_eff), ddm_analytic.mean_decision_time(v, a_eff, z_eff)


@pytest.mark.parametrize("kind", ["par2", "seq2"])
def test_two_stage_choices_and_times_match_the_wiener_closed_forms(kind):
    """With a constant bound both walks are independent Wiener exits: the four choice
    probabilities are the products of the absorption probabilities (high walk, then the low
    walk picked by it), and for seq2 the decision time is T_h + T_l, so its mean is the sum of
    the mean exit times; for par2 it is max(T_h, T_l), at least either mean. A swapped low
    walker, a wrong coin or a missing stage moves these beyond the bounds."""
    vh, vl1, vl2, a, zh, zl1, zl2 = 0.9, 0.6, -0.8, 1.0, 0.55, 0.4, 0.6
    n = 40000
    r, c, _, rh, rl = E.two_stage(kind, [vh], [vl1], [vl2], [zh], [zl1], [zl2], [0.0], [999.0], [1.0], n_samples=n,
                                  bparams={"a": [a]}, rng=np.random.default_rng(21 if kind == "par2" else 22))
    ph, th = _siegmund(vh, a, zh)
    p1, t1 = _siegmund(vl1, a, zl1)
    p2, t2 = _siegmund(vl2, a, zl2)
    expect = [(1 - ph) * (1 - p1), (1 - ph) * p1, ph * (1 - p2), ph * p2]
    for j in range(4):
        pj = np.mean(c == j)
        assert abs(pj - expect[j]) < 4 * math.sqrt(expect[j] * (1 - expect[j]) / n) + 1e-3, (j, pj, expect[j])
    tl = (1 - ph) * t1 + ph * t2
    if kind == "seq2":
        assert abs(r.mean() - (th + tl)) < 4 * r.std() / math.sqrt(n) + 0.004
        assert rh is None and rl is None
    else:
        assert r.mean() > max(th, tl) and r.mean() < th + tl
        assert abs(rh.mean() - th) < 4 * rh.std() / math.sqrt(n) + 0.003
        assert np.array_equal(r, np.maximum(rh, rl))


def test_two_stage_high_walk_is_the_ddm_walk_on_the_same_draws():
    """Replay: the high-level walk consumes the first draws of each row exactly as the
    single-stage walk with the same float start (the Ornstein form at zero leak) does."""
    rng = np.random.default_rng(23)
    K, N, L = 3, 200, 4001
    vh = rng.uniform(-1, 1, K).astype(F32)
    zh = rng.uniform(0.3, 0.7, K).astype(F32)
    a = rng.uniform(0.8, 1.5, K).astype(F32)
    nz = rng.standard_normal((K * N, L)).astype(F32)
    uu = rng.uniform(size=(K * N, 4)).astype(F32)
    ones, zeros = np.ones(K, F32), np.zeros(K, F32)
    r0, _, s0 = E.ddm_flexbound(vh, a, zh, zeros, np.full(K, 999, F32), ones, n_samples=N, delta_t=DT, g=zeros,
                                normals=nz)
    _, _, _, rh, _ = E.two_stage("par2", vh, vh, vh, zh, zh, zh, zeros, np.full(K, 999, F32), ones, n_samples=N,
                                 bparams={"a": a}, normals=nz, uniforms=uu)
    assert np.array_equal(rh, r0)


def test_seq2_stops_at_max_t_with_the_prior_low_choice():
    """A walk that never leaves wide bounds reaches max_t: seq2 then takes the high choice by
    the coin rule at the walk's position and the low choice with probability zl, and reports
    t_particle (+ t0) as rt."""
    n = 4000
    r, c, _, _, _ = E.two_stage("seq2", [0.0], [0.0], [0.0], [0.5], [0.25], [0.8], [0.1], [999.0], [0.05],
                                n_samples=n, max_t=2.0, bparams={"a": [50.0]}, rng=np.random.default_rng(24))
    assert np.all(r > 2.0)
    hi = c >= 2
    assert abs(np.mean(hi) - 0.5) < 4 * math.sqrt(0.25 / n)  # y stays near 0: ratio ~ 1/2
    assert abs(np.mean(c[~hi] == 1) - 0.25) < 4 * math.sqrt(0.25 * 0.75 / max(1, (~hi).sum()))
    assert abs(np.mean(c[hi] == 3) - 0.8) < 4 * math.sqrt(0.8 * 0.2 / max(1, hi.sum()))


def test_seq2_low_start_outside_the_bound_takes_the_prior_choice_without_walking():
    """Low start -b + 2 zl b outside [-b, b] (zl = 1.2 above, -0.2 below): seq2 takes the low
    choice by u < zl (always +1, never +1) and does not walk, so rt and the step count are the
    high walk's, which the single-stage walk on the same draws gives."""
    rng = np.random.default_rng(25)
    K, N, L = 2, 300, 4001
    vh = np.array([0.5, -0.4], F32)
    zh = np.array([0.5, 0.6], F32)
    a = np.array([1.0, 0.8], F32)
    zl1 = np.array([1.2, -0.2], F32)
    zl2 = np.array([-0.2, 1.2], F32)
    nz = rng.standard_normal((K * N, L)).astype(F32)
    uu = rng.uniform(size=(K * N, 4)).astype(F32)
    t0, zeros, ones = np.array([0.2, 0.3], F32), np.zeros(K, F32), np.ones(K, F32)
    r, c, steps, _, _ = E.two_stage("seq2", vh, vh, vh, zh, zl1, zl2, t0, np.full(K, 999, F32), ones, n_samples=N,
                                    bparams={"a": a}, normals=nz, uniforms=uu)
    r0, c0, s0 = E.ddm_flexbound(vh, a, zh, t0, np.full(K, 999, F32), ones, n_samples=N, delta_t=DT, g=zeros,
                                 normals=nz)
    assert np.array_equal(steps, s0) and np.array_equal(r, r0)
    hi = np.repeat(np.arange(K), N)
    zl_used = np.where(c >= 2, zl2[hi], zl1[hi])
    assert np.array_equal(c % 2 == 1, zl_used > 1)
    assert np.array_equal(c >= 2, c0 == 1)  # the high coin is decided by the exit side


# ----------------------------------------------------------------------------- legacy DDM
@pytest.mark.parametrize("v,a,z", [(1.0, 1.5, 0.5), (-0.7, 1.0, 0.35)])
def test_legacy_ddm_matches_wiener_closed_form(v, a, z):
    """Bounds 0 and a from z a: the Wiener absorption probability and mean exit time of the
    interval (0, a) (both bounds moved out by beta sqrt(dt))."""
    n = 20000
    rts, ch, _ = E.ddm_flexbound([v], [a], [z], [0.0], [999.0], [1.0], n_samples=n, delta_t=DT, legacy=True,
                                 rng=np.random.default_rng(31))
    sh = ddm_analytic.SIEGMUND_BETA * SQDT
    # ddm_analytic works on (-A, A) from (2z' - 1) A: width a + 2 sh, start z a + sh from the lower bound
    A = (a + 2 * sh) / 2
    z_eff = (z * a + sh) / (2 * A)
    p = ddm_analytic.p_upper(v, A, z_eff)
    assert abs(np.mean(ch == 1) - p) < 4 * math.sqrt(p * (1 - p) / n)
    mt = ddm_analytic.mean_decision_time(v, A, z_eff)
    assert abs(rts.mean() - mt) < 4 * rts.std() / math.sqrt(n) + 0.003


def test_legacy_ddm_is_the_symmetric_walk_shifted_by_half_the_bound():
    """With the same draws, the [0, a] walk from z a and the [-a/2, a/2] walk from (z - 1/2) a
    differ by the constant a/2 only; with a and z exact in binary and a short walk, float
    rounding leaves the two crossing steps and choices equal."""
    rng = np.random.default_rng(32)
    K, N, L = 2, 300, 2001
    nz = rng.standard_normal((K * N, L)).astype(F32)
    v = np.array([0.5, -1.0], F32)
    z = np.array([0.5, 0.25], F32)
    a = np.array([2.0, 1.0], F32)
    t = np.zeros(K, F32)
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 legacy=True, normals=nz)
    r1, c1, s1 = E.ddm_flexbound(v, a / 2, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=DT,
                                 normals=nz)
    assert np.mean(s0 == s1) > 0.99 and np.mean(c0 == c1) > 0.99


@pytest.mark.parametrize("bname", ["constant", "angle"])
def test_oracle_trajectory_of_trial0_sample0_is_its_walk(bname):
    # the recorded positions are walk (0, 0) itself: the start in row 0 (ddm_start32), one row
    # per step it took, every row but the last inside the bounds of its step index, the last
    # one outside (or the clock past max_t), its sign the choice; -999 after it
    K, N, L_ = 3, 40, 257
    rng = np.random.default_rng(8)
    v = np.array([0.8, -0.5, 1.2], F32)
    a = np.array([1.1, 0.9, 1.4], F32)
    z = np.array([0.4, 0.6, 0.5], F32)
    bp = {"a": a, "theta": np.array([0.3, 0.1, 0.2], F32)} if bname == "angle" else {"a": a}
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    r, c, steps, tr = E.ddm_flexbound(v, a, z, np.zeros(K, F32), np.full(K, 999, F32), np.ones(K, F32),
                                      n_samples=N, delta_t=DT, boundary=bname, bparams=bp, normals=normals,
                                      record0=True)
    n = int(steps[0])
    B = E.boundary_table(bname, bp, K, 20.0, DT)[0]
    assert tr.shape == (E.n_steps(20.0, DT), 1)
    assert tr[0, 0] == E.ddm_start32(z[:1], B[:1])[0]
    assert np.all(tr[: n + 1, 0] != -999.0) and np.all(tr[n + 1:, 0] == -999.0)
    assert np.all(np.abs(tr[:n, 0]) <= B[:n])
    assert abs(tr[n, 0]) > B[n]
    assert np.sign(tr[n, 0]) == c[0]
    sst = E.sqrt_dt32(DT)
    for i in range(n):  # step i: y + ((v dt) + (sqrt_st g_i)) in float32
        inc = F32(F32(v[0] * F32(DT)) + F32(sst * normals[0, i % L_]))
        assert tr[i + 1, 0] == F32(tr[i, 0] + inc)


@pytest.mark.parametrize("model", ["race", "lca"])
def test_oracle_accumulator_trajectory_of_trial0_sample0(model):
    K, P, N, L_ = 2, 3, 30, 311
    rng = np.random.default_rng(9)
    v = rng.uniform(0.5, 2.0, (K, P)).astype(F32)
    z = rng.uniform(0.0, 0.4, (K, P)).astype(F32)
    a = np.array([1.5, 2.0], F32)
    normals = rng.standard_normal((K * N, L_)).astype(F32)
    kw = dict(g=np.full(K, 0.3, F32), b=np.full(K, 0.2, F32)) if model == "lca" else {}
    r, c, steps, tr = E.accumulators(model, v, z, np.zeros(K, F32), np.ones((K, 1), F32), np.full(K, 999, F32),
                                     n_samples=N, delta_t=DT, bparams={"a": a}, normals=normals, record0=True, **kw)
    n = int(steps[0])
    assert tr.shape == (E.n_steps(20.0, DT), P)
    assert np.array_equal(tr[0], (z[0] * a[0]).astype(F32))
    assert np.all(tr[: n + 1] != -999.0) and np.all(tr[n + 1:] == -999.0)
    assert np.all(tr[:n].max(axis=1) <= a[0]) and tr[n].max() > a[0]
    assert int(np.argmax(tr[n])) == c[0]
    assert np.all(tr[: n + 1] >= 0.0)
