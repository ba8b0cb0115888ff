"""Float32 NumPy transcription of the reference's sequential EM loops (test oracle).

Follows, operation by operation in float32:
  ddm_flexbound  src/cssm/ddm_models.pyx:696-744  (start :709, walk :719-733,
                 smooth_unif :735-739, rt :741, sign :742, deadline :744)
  ddm            src/cssm/ddm_models.pyx:453-482 (legacy bounds 0 and a)
  full_ddm       src/cssm/ddm_models.pyx:2284-2358 (variability of v, z, t)
  ornstein_uhlenbeck  src/cssm/ornstein_models.pyx:156-198 (leaky drift)
  levy_flexbound src/cssm/levy_models.pyx:160-220 (alpha-stable increments, CMS
                 sampler _utils.pyx:369-470)
  race_model     src/cssm/race_models.pyx:417-464
  lca            src/cssm/race_models.pyx:825-879
  racing_diffusion_model  src/cssm/race_models.pyx:1150-1280 (no floor, U(0, A) starts,
                 in-step first-past-the-post)
  ddm_flexbound_par2 / ddm_flexbound_seq2  src/cssm/parallel_models.pyx:364-560,
                 src/cssm/sequential_models.pyx:377-640 (two-stage hierarchical DDMs)
  lba_vanilla / lba_angle  src/cssm/lba_models.pyx:101-115, 212-224
with helpers compute_smooth_unif / enforce_deadline / compute_deadline_tmp
(src/cssm/_utils.pyx:668-718) and boundary functions
(ssms/basic_simulators/boundary_functions.py:14-122).

Vectorised across trajectories (one array lane per trajectory), looping over
time steps. Normal increments come from either
  * replay: float32[n_total, stride]; draw d of trajectory i is
    normals[i, d % stride] with d = step * n_particles + particle, or
  * a NumPy Generator: fresh float32 standard normals every step.
Trajectory i belongs to set k = i // n_samples (set-major layout).

Pinning: the DDM walk is pinned by the closed-form first-passage
probability / mean decision time (tests/test_oracle_em.py, Siegmund-corrected)
and the noise-free crossing step; race and LCA by the one-accumulator
first-passage closed form, the LCA(g=b=0) == race identity and the
noise-free crossing step; LBA by its closed-form choice probability.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def time_grid(max_t: float, delta_t: float) -> np.ndarray:
    """t_s = np.arange(0, max_t + dt, dt).astype(float32), _utils.pyx:636-640."""
    dt = F32(delta_t)
    stop = float(F32(F32(max_t) + dt))
    return np.arange(0, stop, float(dt)).astype(F32)


def n_steps(max_t: float, delta_t: float) -> int:
    """Length of t_s: the boundary array the sequential loop indexes (ddm_models.pyx:668)."""
    return int(time_grid(max_t, delta_t).shape[0])


def b_constant(t, a):
    return np.full_like(t, a)


def b_angle(t, a, theta):
    return a + np.multiply(t, (-np.sin(theta) / np.cos(theta)))


def b_weibull(t, a, alpha, beta):
    return a * np.exp(-np.power(np.divide(t, beta), alpha))


BOUNDARIES = {"constant": (b_constant, ["a"]), "angle": (b_angle, ["a", "theta"]),
              "weibull_cdf": (b_weibull, ["a", "alpha", "beta"])}


def boundary_table(name: str, params: dict, n_trials: int, max_t: float, delta_t: float) -> np.ndarray:
    """Per-trial float32 boundary over t_s (compute_boundary, _utils.pyx:655-666)."""
    fun, names = BOUNDARIES[name]
    t_s = time_grid(max_t, delta_t)
    out = np.empty((n_trials, t_s.shape[0]), dtype=F32)
    for k in range(n_trials):
        kw = {p: F32(np.asarray(params[p], F32).reshape(n_trials, -1)[k, 0]) for p in names}
        out[k] = np.asarray(fun(t_s, **kw)).astype(F32)
    return out


def sqrt_dt32(delta_t: float) -> np.float32:
    return F32(np.sqrt(np.float64(F32(delta_t))))


def smooth_unif32(t: np.ndarray, dl_tmp: np.ndarray, dt: np.float32, u: np.ndarray) -> np.ndarray:
    """compute_smooth_unif (_utils.pyx:668-695): float arguments, but the 0.5 literals are C
    doubles, so each branch is formed in double and rounded to float on return."""
    u64 = np.asarray(u, F32).astype(np.float64)
    half = (u64 * 0.5 * np.float64(dt)).astype(F32)
    mid = ((0.5 - u64) * np.float64(dt)).astype(F32)
    return np.where(t == F32(0.0), half, np.where(t < dl_tmp, mid, F32(0.0))).astype(F32)


def ddm_start32(z, b0, double_literal=False):
    """Start point -b0 + 2 z b0 of the two-bound diffusions, in the reference's C precision.

    ddm_flexbound (src/cssm/ddm_models.pyx:709) writes `(-1) * b0 + (z * 2 * b0)` with an
    int 2 and float operands, so C evaluates it in float: fl(-b0 + fl(fl(2 z) b0)), two
    roundings (ornstein_models.pyx:165, levy_models.pyx:183 and the two-stage models use the
    same form). full_ddm (:2305) and the parallel branch (:798) write `z * 2.0 * b0` with a
    double literal: the sum is formed in double and rounded to float once."""
    z = np.asarray(z, F32)
    b0 = np.asarray(b0, F32)
    if double_literal:
        return ((-b0).astype(np.float64) + (z.astype(np.float64) * 2.0) * b0.astype(np.float64)).astype(F32)
    return ((-b0) + ((z * F32(2)).astype(F32) * b0).astype(F32)).astype(F32)


def stable_cms(u, w, alpha):
    """Symmetric alpha-stable variates by Chambers–Mallows–Stuck, as draw_random_stable
    (src/cssm/_utils.pyx:369-470) forms them in float32 from u ~ U(-pi/2 + 1e-7, pi/2 - 1e-7)
    and w ~ Exp(1) floored at 1e-10: alpha == 1 gives tan(u); otherwise
    sin(alpha u) / |cos u|^(1/alpha) * (|cos((1 - alpha) u)| / w)^((1 - alpha)/alpha) with both
    cosines floored at 1e-10 in magnitude; clipped to +-1e10, NaN -> 0. `alpha` may be a
    per-variate array (one set's alpha per lane)."""
    u = np.asarray(u, F32)
    w = np.maximum(np.asarray(w, F32), F32(1e-10))
    alpha = np.broadcast_to(np.asarray(alpha, F32), u.shape)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        alpha_inv = (F32(1.0) / alpha).astype(F32)
        scale = ((F32(1.0) - alpha) / alpha).astype(F32)
        cu = np.cos(u)
        cu = (np.sign(cu) * np.maximum(np.abs(cu), F32(1e-10))).astype(F32)
        ca = np.cos((u * (F32(1.0) - alpha)).astype(F32))
        ca = (np.sign(ca) * np.maximum(np.abs(ca), F32(1e-10))).astype(F32)
        r = ((np.sin((alpha * u).astype(F32)) / (np.abs(cu) ** alpha_inv)).astype(F32)
             * ((np.abs(ca) / w).astype(F32) ** scale).astype(F32)).astype(F32)
        r = np.where(alpha == F32(1.0), np.tan(u).astype(F32), r)
    r = np.clip(r, F32(-1e10), F32(1e10))
    return np.nan_to_num(r, nan=0.0, posinf=1e10, neginf=-1e10).astype(F32)


def draw_stable(rng, n, alpha):
    """n variates from a NumPy Generator, uniforms then exponentials (_utils.pyx:425-433)."""
    u = rng.uniform(-np.pi / 2 + 1e-7, np.pi / 2 - 1e-7, n).astype(F32)
    w = rng.exponential(1.0, n).astype(F32)
    return stable_cms(u, w, alpha)


def levy_scale32(s, alpha, delta_t):
    """delta_t_alpha = s * pow(delta_t, 1.0 / alpha) (levy_models.pyx:166): float s times a
    double pow, rounded to float once."""
    s64 = np.asarray(s, F32).astype(np.float64)
    a64 = np.asarray(alpha, F32).astype(np.float64)
    return (s64 * np.power(np.float64(F32(delta_t)), 1.0 / a64)).astype(F32)


class _Trajectory:
    """Positions of lane 0 (trial 0, sample 0): the start in row 0, then the position after the
    step that took its step index to ix in row ix; -999 elsewhere (the reference's traj buffer)."""

    def __init__(self, on, n_rows, n_particles, start):
        self.on = bool(on)
        self.rows = np.full((n_rows, n_particles), F32(-999.0), F32) if self.on else None
        if self.on:
            self.rows[0] = np.asarray(start, F32).reshape(-1)

    def after_step(self, moved, ix, pos0):
        if self.on and moved.size and moved[0] == 0 and ix[0] < self.rows.shape[0]:
            self.rows[ix[0]] = np.asarray(pos0, F32).reshape(-1)


class _Normals:
    def __init__(self, normals=None, rng=None, n_lanes=0):
        self.normals = None if normals is None else np.asarray(normals, F32)
        self.rng = rng
        if self.normals is not None:
            self.stride = self.normals.shape[1]

    def draw(self, lanes: np.ndarray, d: np.ndarray) -> np.ndarray:
        if self.normals is not None:
            return self.normals[lanes, d % self.stride]
        return self.rng.standard_normal(lanes.shape[0], dtype=F32)


def _finish(rts, t_el, t0, dl_tmp, deadline, dt, smooth, uniforms):
    su = smooth_unif32(t_el, dl_tmp, dt, uniforms) if smooth else np.zeros_like(t_el)
    rt = ((t_el + t0).astype(F32) + su).astype(F32)
    omit = (rt >= deadline) | (deadline <= F32(0.0))
    rt[omit] = F32(-999.0)
    return rt


def ddm_flexbound(v, a, z, t, deadline, s, *, n_samples, max_t=20.0, delta_t=0.001, boundary="constant",
                  bparams=None, smooth_unif=False, normals=None, uniforms=None, rng=None, table=None,
                  sz=None, sv=None, st=None, g=None, alpha=None, legacy=False, record0=False):
    """Returns (rts float32[n_total], choices int32[n_total], steps int64[n_total]), set-major.
    With record0: a 4th value, the trajectory of trial 0, sample 0 as the sequential loop
    records it (ddm_models.pyx:714-716, :726-728): float32[len(t_s), 1], the start in row 0,
    the position after step ix in row ix, -999 after the walk's end.

    With sz/sv/st: full_ddm (src/cssm/ddm_models.pyx:2284-2358): per trajectory the drift
    increment (v + sv gv) dt (float), t_tmp = t + 2 (u_t - 0.5) st and the start shift
    2 (u_z - 0.5) sz (both formed in double: 0.5 is a C double), rt = t_particle + t_tmp +
    smooth_u, deadline_tmp from the set's t. Replay: normals[i, 0] is gv and step d reads
    normals[i, (1 + d) % stride]; uniforms[i] = (u_t, u_z, u_smooth).
    With g: ornstein_uhlenbeck (src/cssm/ornstein_models.pyx:156-198): float start
    -b0 + (2z) b0 (the literal 2 is an int) and step ((v - g y) dt) + sqrt_st gm.
    With alpha: levy_flexbound (src/cssm/levy_models.pyx:160-220): the same float start, step
    (v dt) + (delta_t_alpha x) with x symmetric alpha-stable (stable_cms) and
    delta_t_alpha = levy_scale32(s, alpha, dt); replay `normals` then hold the stable variates.
    With legacy=True: ddm (src/cssm/ddm_models.pyx:453-482): bounds 0 and a (inclusive), float
    start z a, the same step; `a` is the constant bound."""
    full = sz is not None or sv is not None or st is not None
    ou = g is not None
    levy = alpha is not None
    v, a, z, t, deadline, s = (np.asarray(x, F32).reshape(-1) for x in (v, a, z, t, deadline, s))
    K = v.shape[0]
    N = int(n_samples)
    dt = F32(delta_t)
    bparams = dict(bparams or {"a": a})
    B = table if table is not None else boundary_table(boundary, bparams, K, max_t, delta_t)
    nst = B.shape[1]
    dl_tmp_k = np.minimum(F32(max_t), (deadline - t).astype(F32)).astype(F32)
    sqrt_st_k = (sqrt_dt32(delta_t) * s).astype(F32)
    vdt_k = (v * dt).astype(F32)
    total = K * N
    lane_set = np.repeat(np.arange(K), N)
    b0 = B[lane_set, 0]
    if levy:
        alpha_k = np.asarray(alpha, F32).reshape(-1)
        sqrt_st_k = levy_scale32(s, alpha_k, delta_t)
    if legacy:
        y = (z[lane_set] * b0).astype(F32)
    elif full:
        y = ddm_start32(z[lane_set], b0, double_literal=True)
    else:
        y = ddm_start32(z[lane_set], b0)
    t_lane = t[lane_set].copy()
    vdt_lane = vdt_k[lane_set].copy()
    draw_off = 0
    if full:
        sz_, sv_, st_ = (np.zeros(K, F32) if x is None else np.asarray(x, F32).reshape(-1) for x in (sz, sv, st))
        if normals is not None:
            gv = np.asarray(normals, F32)[:, 0]
            uu = np.asarray(uniforms, F32).reshape(total, 3)
            u_t, u_z = uu[:, 0], uu[:, 1]
            draw_off = 1
        else:
            gv = rng.standard_normal(total, dtype=F32)
            u_t = rng.uniform(0.0, 1.0, total).astype(F32)
            u_z = rng.uniform(0.0, 1.0, total).astype(F32)
        vdt_lane = ((v[lane_set] + (sv_[lane_set] * gv).astype(F32)).astype(F32) * dt).astype(F32)
        t_lane = (t[lane_set].astype(np.float64) + (2.0 * (u_t.astype(np.float64) - 0.5))
                  * st_[lane_set].astype(np.float64)).astype(F32)
        y = (y.astype(np.float64) + (2.0 * (u_z.astype(np.float64) - 0.5)) * sz_[lane_set].astype(np.float64)).astype(F32)
    if ou:
        g_k = np.asarray(g, F32).reshape(-1)
    tt = np.zeros(total, F32)
    ix = np.zeros(total, np.int64)
    src = _Normals(normals, rng)
    traj = _Trajectory(record0 and total > 0, nst, 1, y[:1])
    active = np.arange(total)
    while active.size:
        ks = lane_set[active]
        ixa = ix[active]
        inb = ixa < nst
        b = B[ks, np.minimum(ixa, nst - 1)]
        ya = y[active]
        lo = F32(0.0) if legacy else -b
        go = inb & (ya >= lo) & (ya <= b) & (tt[active] <= dl_tmp_k[ks])
        active = active[go]
        if not active.size:
            break
        ks = lane_set[active]
        if levy and normals is None:
            gm = draw_stable(rng, active.size, alpha_k[ks])
        else:
            gm = src.draw(active, ix[active] + draw_off)
        noise = (sqrt_st_k[ks] * gm).astype(F32)
        if ou:
            ya = y[active]
            drift = ((v[ks] - (g_k[ks] * ya).astype(F32)).astype(F32) * dt).astype(F32)
            inc = (drift + noise).astype(F32)
        else:
            inc = (vdt_lane[active] + noise).astype(F32)
        y[active] = (y[active] + inc).astype(F32)
        tt[active] = (tt[active] + dt).astype(F32)
        ix[active] += 1
        traj.after_step(active, ix, y[:1])
    if uniforms is None:
        u = np.zeros(total, F32)
    elif full:
        u = np.asarray(uniforms, F32).reshape(total, 3)[:, 2]
    else:
        u = np.asarray(uniforms, F32).reshape(-1)
    if smooth_unif and uniforms is None:
        u = (rng.uniform(0.0, 1.0, total)).astype(F32)
    rts = _finish(None, tt, t_lane, dl_tmp_k[lane_set], deadline[lane_set], dt, smooth_unif, u)
    choices = ((y > 0).astype(np.int32) - (y < 0).astype(np.int32)).astype(np.int32)
    return (rts, choices, ix, traj.rows) if record0 else (rts, choices, ix)


def accumulators(model, v, z, t, s, deadline, *, n_samples, max_t=20.0, delta_t=0.001, boundary="constant",
                 bparams=None, g=None, b=None, smooth_unif=False, normals=None, uniforms=None, rng=None,
                 table=None, record0=False):
    """race_model (model='race') or lca (model='lca'); returns (rts, choices, steps), set-major.
    With record0: a 4th value, the trajectory of trial 0, sample 0 (race_models.pyx:433-452,
    :841-868): float32[len(t_s), P], the starts in row 0, positions after step ix in row ix."""
    v = np.asarray(v, F32)
    K, P = v.shape
    z = np.broadcast_to(np.asarray(z, F32), (K, P))
    s = np.broadcast_to(np.asarray(s, F32).reshape(K, -1), (K, P))
    t = np.asarray(t, F32).reshape(K, -1)[:, 0]
    deadline = np.asarray(deadline, F32).reshape(-1)
    N = int(n_samples)
    dt = F32(delta_t)
    bparams = dict(bparams or {})
    B = table if table is not None else boundary_table(boundary, bparams, K, max_t, delta_t)
    nst = B.shape[1]
    dl_tmp_k = np.minimum(F32(max_t), (deadline - t).astype(F32)).astype(F32)
    sst = (sqrt_dt32(delta_t) * s).astype(F32)
    vdt = (v * dt).astype(F32)
    gk = None if g is None else np.asarray(g, F32).reshape(K, -1)[:, 0]
    bk = None if b is None else np.asarray(b, F32).reshape(K, -1)[:, 0]
    total = K * N
    lane_set = np.repeat(np.arange(K), N)
    x = (z[lane_set] * B[lane_set, 0][:, None]).astype(F32)
    tt = np.zeros(total, F32)
    ix = np.zeros(total, np.int64)
    src = _Normals(normals, rng)
    traj = _Trajectory(record0 and total > 0, nst, P, x[0])
    active = np.arange(total)
    while active.size:
        ks = lane_set[active]
        ixa = ix[active]
        bnd = B[ks, np.minimum(ixa, nst - 1)]
        crossed = np.any(x[active] > bnd[:, None], axis=1)
        go = (ixa < nst) & ~crossed & (tt[active] <= dl_tmp_k[ks])
        active = active[go]
        if not active.size:
            break
        ks = lane_set[active]
        xa = x[active]
        noise = np.empty_like(xa)
        d0 = ix[active] * P
        for j in range(P):
            noise[:, j] = src.draw(active, d0 + j)
        if model == "race":
            xn = (xa + (vdt[ks] + (sst[ks] * noise).astype(F32)).astype(F32)).astype(F32)
        else:
            tot = np.zeros(active.size, F32)
            for j in range(P):  # csum: sequential float32 sum
                tot = (tot + xa[:, j]).astype(F32)
            red = ((-xa) + tot[:, None]).astype(F32)
            drift = ((v[ks] - (gk[ks, None] * xa).astype(F32)).astype(F32) - (bk[ks, None] * red).astype(F32)).astype(F32)
            drift = (drift * dt).astype(F32)
            xn = (xa + (drift + (sst[ks] * noise).astype(F32)).astype(F32)).astype(F32)
        x[active] = np.maximum(F32(0.0), xn)
        tt[active] = (tt[active] + dt).astype(F32)
        ix[active] += 1
        traj.after_step(active, ix, x[0])
    u = np.zeros(total, F32) if uniforms is None else np.asarray(uniforms, F32).reshape(-1)
    if smooth_unif and uniforms is None:
        u = (rng.uniform(0.0, 1.0, total)).astype(F32)
    rts = _finish(None, tt, t[lane_set], dl_tmp_k[lane_set], deadline[lane_set], dt, smooth_unif, u)
    choices = np.argmax(x, axis=1).astype(np.int32)
    return (rts, choices, ix, traj.rows) if record0 else (rts, choices, ix)


def racing_diffusion(v, A, b, t, s, deadline, *, n_samples, max_t=20.0, delta_t=0.001, smooth_unif=False,
                     normals=None, uniforms=None, rng=None, record0=False):
    """racing_diffusion_model, sequential path (src/cssm/race_models.pyx:1150-1280):
    accumulators start at u_j A (:1207-1208); each step moves them in order by
    (v_j dt) + (sqrt_st g) and the first to reach b wins at once, before the rest move and
    before t advances (:1218-1237); rt = t + t0 + smooth_u, and past the deadline or without
    a winner (-999, -1) (:1248-1255). Replay: normals[i, (step * P + j) % stride];
    uniforms[i] = (u_0 .. u_{P-1}, u_smooth). Returns (rts, choices, steps) set-major, steps
    counting the winning step; with record0 also the trajectory of trial 0, sample 0 (starts in
    row 0, positions after every step without a winner, race_models.pyx:1223-1254)."""
    v = np.asarray(v, F32)
    K, P = v.shape
    A = np.asarray(A, F32).reshape(K, -1)[:, 0]
    b = np.asarray(b, F32).reshape(K, -1)[:, 0]
    t = np.asarray(t, F32).reshape(K, -1)[:, 0]
    s = np.asarray(s, F32).reshape(K, -1)[:, 0]
    deadline = np.asarray(deadline, F32).reshape(-1)
    N = int(n_samples)
    dt = F32(delta_t)
    total = K * N
    lane_set = np.repeat(np.arange(K), N)
    dl_tmp_k = np.minimum(F32(max_t), (deadline - t).astype(F32)).astype(F32)
    sst_k = (sqrt_dt32(delta_t) * s).astype(F32)
    vdt = (v * dt).astype(F32)
    if uniforms is not None:
        uu = np.asarray(uniforms, F32).reshape(total, P + 1)
    else:
        uu = rng.uniform(0.0, 1.0, (total, P + 1)).astype(F32)
    x = (uu[:, :P] * A[lane_set][:, None]).astype(F32)
    tt = np.zeros(total, F32)
    ix = np.zeros(total, np.int64)
    win = np.full(total, -1, np.int64)
    src = _Normals(normals, rng)
    traj = _Trajectory(record0 and total > 0, n_steps(max_t, delta_t), P, x[0])
    active = np.arange(total)
    while active.size:
        ks = lane_set[active]
        active = active[tt[active] <= dl_tmp_k[ks]]
        if not active.size:
            break
        ks = lane_set[active]
        found = np.zeros(active.size, bool)
        for j in range(P):  # in order; a lane that has a winner moves no further accumulator
            mv = ~found
            lanes = active[mv]
            gm = src.draw(lanes, ix[lanes] * P + j)
            x[lanes, j] = (x[lanes, j] + (vdt[ks[mv], j] + (sst_k[ks[mv]] * gm).astype(F32)).astype(F32)).astype(F32)
            hit = x[lanes, j] >= b[ks[mv]]
            win[lanes[hit]] = j
            idx = np.flatnonzero(mv)[hit]
            found[idx] = True
        cont = active[~found]
        tt[cont] = (tt[cont] + dt).astype(F32)
        ix[cont] += 1
        traj.after_step(cont, ix, x[0])
        active = cont
    su = smooth_unif32(tt, dl_tmp_k[lane_set], dt, uu[:, P]) if smooth_unif else np.zeros(total, F32)
    rt = ((tt + t[lane_set]).astype(F32) + su).astype(F32)
    dl = deadline[lane_set]
    rt[(rt >= dl) | (dl <= F32(0.0))] = F32(-999.0)
    bad = (rt == F32(-999.0)) | (win < 0)
    rt[bad] = F32(-999.0)
    ch = np.where(bad, -1, win).astype(np.int32)
    return (rt, ch, ix + (win >= 0), traj.rows) if record0 else (rt, ch, ix + (win >= 0))


def _walk(y, tt, ix, moving, vdt, sst, B, lane_set, dl_tmp, src, draw_base, dt):
    """Advance the lanes in `moving` as one DDM walk between -B and B (inclusive), step
    (v dt) + (sqrt_st g), until each leaves, passes dl_tmp or reaches the end of t_s.
    Draw d of a lane's walk reads index draw_base + d. Returns the steps each lane took."""
    nst = B.shape[1]
    steps = np.zeros(y.shape[0], np.int64)
    active = np.flatnonzero(moving)
    while active.size:
        ks = lane_set[active]
        ixa = ix[active]
        b = B[ks, np.minimum(ixa, nst - 1)]
        ya = y[active]
        active = active[(ixa < nst) & (ya >= -b) & (ya <= b) & (tt[active] <= dl_tmp[ks])]
        if not active.size:
            break
        gm = src.draw(active, draw_base[active] + steps[active])
        inc = (vdt[active] + (sst[lane_set[active]] * gm).astype(F32)).astype(F32)
        y[active] = (y[active] + inc).astype(F32)
        tt[active] = (tt[active] + dt).astype(F32)
        ix[active] += 1
        steps[active] += 1
    return steps


def _coin(u, y, b):
    """The choice rule of the two-stage models: with b <= 0 a fair coin (u <= 0.5), else
    u <= (y + b) / (2 b) in float32 (parallel_models.pyx:474-479, :547-552)."""
    ratio = ((y + b).astype(F32) / (F32(2) * b).astype(F32)).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(b <= F32(0.0), u <= F32(0.5), u <= ratio)


def two_stage(kind, vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, *, n_samples, max_t=20.0, delta_t=0.001,
              boundary="constant", bparams=None, smooth_unif=False, normals=None, uniforms=None, rng=None,
              table=None):
    """Hierarchical two-choice DDMs, sequential paths: kind "par2" = ddm_flexbound_par2
    (src/cssm/parallel_models.pyx:364-560), kind "seq2" = ddm_flexbound_seq2
    (src/cssm/sequential_models.pyx:377-640). A high-level walk (vh, zh) picks the high choice
    by the coin rule at its exit; a low-level walk (vl1, zl1 after high choice 0, vl2, zl2 after
    2) then picks +1. par2: the low walk starts at t = 0 from b(0) and rt = max(t_h, t_l);
    seq2: it starts where the high walk ended, from b(ix), and rt is its end time; seq2 stops
    at the high walk when t_h >= max_t (low choice u <= zl) and takes the low choice u < zl
    without walking when the low start lies outside the bounds.
    Replay: normals[i] holds the high walk's draws then the low walk's; uniforms[i] =
    (u_high, u_low_prior, u_smooth, u_low). Returns (rts, choices, steps, rts_high, rts_low),
    the last two for par2 only (None for seq2)."""
    vh, vl1, vl2, zh, zl1, zl2, t, deadline, s = (np.asarray(x, F32).reshape(-1)
                                                  for x in (vh, vl1, vl2, zh, zl1, zl2, t, deadline, s))
    K = vh.shape[0]
    N = int(n_samples)
    dt = F32(delta_t)
    bparams = dict(bparams or {})
    B = table if table is not None else boundary_table(boundary, bparams, K, max_t, delta_t)
    nst = B.shape[1]
    total = K * N
    lane_set = np.repeat(np.arange(K), N)
    dl_tmp_k = np.minimum(F32(max_t), (deadline - t).astype(F32)).astype(F32)
    sst_k = (sqrt_dt32(delta_t) * s).astype(F32)
    uu = (np.asarray(uniforms, F32).reshape(total, 4) if uniforms is not None
          else rng.uniform(0.0, 1.0, (total, 4)).astype(F32))
    src = _Normals(normals, rng)

    def bound_at(ix):
        return np.where(ix < nst, B[lane_set, np.minimum(ix, nst - 1)], F32(0.0)).astype(F32)

    b0 = B[lane_set, 0]
    y = ((-b0) + ((zh[lane_set] * F32(2)).astype(F32) * b0).astype(F32)).astype(F32)
    tt = np.zeros(total, F32)
    ix = np.zeros(total, np.int64)
    nh = _walk(y, tt, ix, np.ones(total, bool), (vh * dt).astype(F32)[lane_set], sst_k, B, lane_set, dl_tmp_k, src,
               np.zeros(total, np.int64), dt)
    t_h = tt.copy()
    hi = _coin(uu[:, 0], y, bound_at(ix))
    choice = np.where(hi, 2, 0).astype(np.int32)
    zl = np.where(hi, zl2[lane_set], zl1[lane_set]).astype(F32)
    vl = np.where(hi, vl2[lane_set], vl1[lane_set]).astype(F32)
    decided = np.zeros(total, bool)
    if kind == "par2":
        yl = ((-b0) + ((zl * F32(2)).astype(F32) * b0).astype(F32)).astype(F32)
        tl = np.zeros(total, F32)
        ixl = np.zeros(total, np.int64)
        nl = _walk(yl, tl, ixl, np.ones(total, bool), (vl * dt).astype(F32), sst_k, B, lane_set, dl_tmp_k, src, nh,
                   dt)
        t_end = np.maximum(t_h, tl)
    elif kind == "seq2":
        at_max = t_h >= F32(max_t)
        choice += (at_max & (uu[:, 1] <= zl)).astype(np.int32)
        decided |= at_max
        bx = bound_at(ix)
        yl = ((-bx) + ((zl * F32(2)).astype(F32) * bx).astype(F32)).astype(F32)
        oob = ~at_max & ((yl >= bx) | (yl <= -bx))
        choice += (oob & (uu[:, 1] < zl)).astype(np.int32)
        decided |= oob
        tl, ixl = t_h.copy(), ix.copy()
        nl = _walk(yl, tl, ixl, ~at_max, (vl * dt).astype(F32), sst_k, B, lane_set, dl_tmp_k, src, nh, dt)
        t_end = tl
    else:
        raise ValueError(f"unknown two-stage kind {kind!r}")
    su = smooth_unif32(t_end, dl_tmp_k[lane_set], dt, uu[:, 2]) if smooth_unif else np.zeros(total, F32)
    if kind == "par2":
        # parallel_models.pyx:538: libc fmax returns a double, so the sum is formed in double
        rt = ((t_end.astype(np.float64) + t[lane_set].astype(np.float64)) + su.astype(np.float64)).astype(F32)
    else:
        rt = ((t_end + t[lane_set]).astype(F32) + su).astype(F32)
    dl = deadline[lane_set]
    rt[(rt >= dl) | (dl <= F32(0.0))] = F32(-999.0)
    choice += (~decided & _coin(uu[:, 3], yl, bound_at(ixl))).astype(np.int32)
    if kind == "par2":
        return rt, choice, nh + nl, (t_h + t[lane_set]).astype(F32), (tl + t[lane_set]).astype(F32)
    return rt, choice, nh + nl, None, None


def lba(v, a, z, t, deadline, sd, *, n_samples, tan_theta=None, uniforms=None, normals=None):
    """lba_vanilla / lba_angle with explicit draws: uniforms, normals float32[n_total, nact]."""
    v = np.asarray(v, F32)
    K, P = v.shape
    N = int(n_samples)
    lane_set = np.repeat(np.arange(K), N)
    a_ = np.asarray(a, F32).reshape(K, -1)[:, 0][lane_set, None]
    z_ = np.asarray(z, F32).reshape(K, -1)[:, 0][lane_set, None]
    sd_ = np.broadcast_to(np.asarray(sd, F32).reshape(K, -1), (K, P))[lane_set]
    zs = (np.asarray(uniforms, F32) * z_).astype(F32)
    vs = np.abs((v[lane_set] + (sd_ * np.asarray(normals, F32)).astype(F32)).astype(F32))
    den = vs if tan_theta is None else (vs + np.asarray(tan_theta, F32).reshape(-1)[lane_set, None]).astype(F32)
    x = ((a_ - zs).astype(F32) / den).astype(F32)
    choices = np.argmin(x, axis=1).astype(np.int32)
    rts = (np.min(x, axis=1) + np.asarray(t, F32).reshape(-1)[lane_set]).astype(F32)
    rts[rts >= np.asarray(deadline, F32).reshape(-1)[lane_set]] = F32(-999.0)
    return rts, choices


def lba64(v, a, z, t, deadline, sd, *, n_samples, tan_theta=None, uniforms=None, normals=None):
    """lba_vanilla / lba_angle in the reference's own precision (src/cssm/lba_models.pyx:101-115,
    :212-224): NumPy float64 throughout, zs = U(0, z) as z u, vs = |v + sd g|, x = (a - zs) / vs
    (angle: / (vs + tan(theta)) with float64 tan of the float32 theta), choice = first argmin,
    rt = min x + t stored to float32 (one rounding), rt >= deadline -> -999. The draws u, g are
    given (float32[n_total, nact]) so that lba() and the kernel can be held against it."""
    f8 = np.float64
    v = np.asarray(v, F32)
    K, P = v.shape
    N = int(n_samples)
    ls = np.repeat(np.arange(K), N)
    a_ = np.asarray(a, F32).reshape(K, -1)[:, 0].astype(f8)[ls, None]
    z_ = np.asarray(z, F32).reshape(K, -1)[:, 0].astype(f8)[ls, None]
    sd_ = np.broadcast_to(np.asarray(sd, F32).reshape(K, -1), (K, P)).astype(f8)[ls]
    zs = np.asarray(uniforms, F32).astype(f8) * z_
    vs = np.abs(v.astype(f8)[ls] + sd_ * np.asarray(normals, F32).astype(f8))
    if tan_theta is not None:
        vs = vs + np.asarray(tan_theta, f8).reshape(-1)[ls, None]
    x = (a_ - zs) / vs
    choices = np.argmin(x, axis=1).astype(np.int32)
    rts = (np.min(x, axis=1) + np.asarray(t, F32).reshape(-1).astype(f8)[ls]).astype(F32)
    rts[rts >= np.asarray(deadline, F32).reshape(-1)[ls]] = F32(-999.0)
    return rts, choices
