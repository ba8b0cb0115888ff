"""Drop-in replacements for the reference's compiled simulators (``cssm.*``).

Each function has the keyword contract of its reference counterpart and
returns ``{'rts': float32[n_samples, n_trials, 1], 'choices':
intc[n_samples, n_trials, 1], 'metadata': {...}}``, so it can be installed as
``model_config["simulator"]`` in the reference's ``Simulator``.  The walk runs
in the gfx950 kernel with the sequential path's semantics (inclusive bounds,
``smooth_unif`` honoured); ``n_threads`` is accepted and validated but the GPU
ignores it.  Seeds drive a counter-based Philox stream, so the same seed gives
the same output on every call (but not the reference's NumPy stream).
"""

from __future__ import annotations

import numpy as np

from . import _device as D
from . import _lib as L
from .boundary_functions import constant, kernel_boundary_code, time_grid
from .em import accumulator_records, boundary_table, ddm_records_device, make_job, two_stage_records, uniform_offsets

MAX_THREADS = 256  # _constants.pxi:10
OMISSION_SENTINEL = -999.0


def _check_threads(n_threads: int) -> None:
    # check_parallel_request, src/cssm/_openmp_status.pyx:177-208
    if n_threads > MAX_THREADS:
        raise ValueError(
            f"n_threads={n_threads} exceeds maximum supported ({MAX_THREADS}). "
            "This is a compile-time limit for thread-local RNG state arrays."
        )


def _seed(random_state) -> int:
    if random_state is None:
        return int(np.random.randint(0, 2**31))
    return int(random_state) & 0xFFFFFFFFFFFFFFFF


def _minimal_meta(name, choices, n_samples, n_trials, boundary_fun_name=None) -> dict:
    # build_minimal_metadata, _utils.pyx:734-755
    meta = {"simulator": name, "possible_choices": choices, "n_samples": n_samples, "n_trials": n_trials}
    if boundary_fun_name is not None:
        meta["boundary_fun_type"] = boundary_fun_name
    return meta


def _full_meta(minimal, params, sim_config, boundary_fun=None, boundary=None, traj=None, boundary_params=None,
               extra=None) -> dict:
    # build_full_metadata, _utils.pyx:757-792
    meta = dict(minimal)
    meta.update(params)
    meta.update(sim_config)
    if boundary_fun is not None:
        meta["boundary_fun_type"] = boundary_fun.__name__
    if boundary is not None:
        meta["boundary"] = boundary
    if traj is not None:
        meta["trajectory"] = traj
    if boundary_params:
        meta.update(boundary_params)
    if extra:
        meta.update(extra)
    return meta


def _boundary_trial0(boundary_fun, boundary_params, max_t, delta_t) -> np.ndarray:
    t_s = time_grid(max_t, delta_t)
    kw = {k: np.asarray(v).reshape(-1)[0] for k, v in boundary_params.items()}
    return np.asarray(boundary_fun(t=t_s, **kw)).astype(np.float32)


_SIDE_STREAMS: dict = {}


def _side_stream():
    """A second stream of the current device (default priority) for the trajectory replay."""
    t = D.torch()
    dev = t.cuda.current_device()
    if dev not in _SIDE_STREAMS:
        _SIDE_STREAMS[dev] = t.cuda.Stream(device=dev)
    return _SIDE_STREAMS[dev]


def _run(model, code, n_particles, rec, n_samples, n_trials, max_t, delta_t, seed, smooth_unif, boundary_fun,
         boundary_params, replay_normals=None, replay_uniforms=None, trajectory=False):
    """One launch over n_trials sets of n_samples. With trajectory=True the replay of trial 0,
    sample 0 (metadata['trajectory']) runs on a side stream beside the launch; the fourth
    value returned is then that float32[len(t_s), n_particles] array, else None."""
    table = None
    if code == L.BOUNDARY_TABLE:
        table = boundary_table(boundary_fun, boundary_params, n_trials, max_t, delta_t)
    job = make_job(
        model, code, n_particles, rec, int(n_samples),
        delta_t=delta_t, max_t=max_t, seed=seed, smooth_unif=smooth_unif,
        out_layout=1, choice_bytes=4, table=table,
        replay_normals=replay_normals, replay_uniforms=replay_uniforms,
    )
    traj_dev = side = None
    if trajectory and n_samples >= 1 and n_trials >= 1:
        t = D.torch()
        inputs_ready = t.cuda.Event()
        inputs_ready.record()  # the job's inputs are enqueued on the current stream
        side = _side_stream()
        side.wait_event(inputs_ready)
        job.launch()
        traj_dev = job.trajectory_async(time_grid(max_t, delta_t).shape[0], stream=side)
    else:
        job.launch()
    total = n_samples * n_trials
    rts, choices = D.to_host(job.rts[:total], job.choices[:total])
    rts = rts.reshape(n_samples, n_trials, 1)
    choices = choices.astype(np.intc, copy=False).reshape(n_samples, n_trials, 1)
    traj = None
    if trajectory:
        if traj_dev is not None:
            side.synchronize()
            traj = traj_dev.cpu().numpy()
        else:  # no walk: every row is -999 (ddm_models.pyx:714-728, race_models.pyx:433-452)
            traj = np.full((time_grid(max_t, delta_t).shape[0], n_particles), -999.0, dtype=np.float32)
    return rts, choices, job, traj


def _diffusion(name, model, v, a, z, t, deadline, s, max_t, delta_t, n_samples, n_trials, boundary_fun,
               boundary_params, random_state, return_option, smooth_unif, n_threads, replay_normals,
               replay_uniforms, extra, params):
    _check_threads(n_threads)
    if return_option not in ("full", "minimal"):
        raise ValueError('return_option must be either "full" or "minimal"')
    boundary_fun = boundary_fun or constant
    boundary_params = dict(boundary_params or {})
    if a is None:  # models whose boundary level arrives only through boundary_params
        if "a" not in boundary_params:
            raise ValueError("boundary_params must provide 'a'")
        a = boundary_params["a"]
    if not boundary_params and boundary_fun is constant:
        boundary_params = {"a": a}
    code = kernel_boundary_code(boundary_fun)
    max_t = float(np.float32(max_t))
    delta_t = float(np.float32(delta_t))
    rec = ddm_records_device(v, a, z, t, deadline, s, max_t, delta_t, code, boundary_params, **extra,
                             stride=L.SET_STRIDE_SINGLE)
    rts, choices, job, traj = _run(model, code, 1, rec, int(n_samples), int(n_trials), max_t, delta_t,
                                   _seed(random_state), smooth_unif, boundary_fun, boundary_params, replay_normals,
                                   replay_uniforms, trajectory=return_option == "full")
    minimal = _minimal_meta(name, [-1, 1], int(n_samples), int(n_trials), boundary_fun.__name__)
    if return_option == "minimal":
        return {"rts": rts, "choices": choices, "metadata": minimal}
    full = _full_meta(
        minimal,
        params,
        {"delta_t": delta_t, "max_t": max_t, "n_threads": n_threads},
        boundary_fun=boundary_fun,
        boundary=_boundary_trial0(boundary_fun, boundary_params, max_t, delta_t) if n_trials > 0 else None,
        traj=traj,
        boundary_params=boundary_params,
    )
    return {"rts": rts, "choices": choices, "metadata": full}


def ddm_flexbound(v, a, z, t, deadline, s, max_t=20, delta_t=0.001, n_samples=20000, n_trials=1,
                  boundary_fun=None, boundary_params={}, random_state=None, return_option="full",
                  smooth_unif=False, n_threads=1, replay_normals=None, replay_uniforms=None, **kwargs):
    """Two-bound DDM with flexible boundary ±b(t) (ddm_models.pyx:589-878)."""
    return _diffusion("ddm_flexbound", L.MODEL_DDM, v, a, z, t, deadline, s, max_t, delta_t, n_samples, n_trials,
                      boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads,
                      replay_normals, replay_uniforms, {},
                      {"v": v, "a": a, "z": z, "t": t, "s": s, "deadline": deadline})


def ddm(v, a, z, t, deadline, s, max_t=20, delta_t=0.001, n_samples=20000, n_trials=10, random_state=None,
        return_option="full", smooth_unif=False, n_threads=1, replay_normals=None, replay_uniforms=None, **kwargs):
    """DDM between the bounds 0 and a, started at z a (ddm_models.pyx:349-582, sequential loop
    :453-482): y += (v dt) + (sqrt(dt) s g) while 0 <= y <= a; choice = sign(y)."""
    n_trials = int(n_trials)
    a32 = np.asarray(a, np.float32).reshape(-1)
    out = _diffusion("ddm", L.MODEL_DDM_LEGACY, v, a32, z, t, deadline, s, max_t, delta_t, n_samples, n_trials,
                     constant, {"a": a32}, random_state, return_option, smooth_unif, n_threads, replay_normals,
                     replay_uniforms, {}, {"v": v, "a": a, "z": z, "t": t, "deadline": deadline, "s": s})
    if return_option == "full":  # the legacy simulator reports no boundary array (ddm_models.pyx:565-573)
        out["metadata"].pop("boundary", None)
    return out


def full_ddm(v, a, z, t, sz, sv, st, deadline, s, delta_t=0.001, max_t=20, n_samples=20000, n_trials=1,
             boundary_fun=None, boundary_params={}, random_state=None, return_option="full", smooth_unif=False,
             n_threads=1, replay_normals=None, replay_uniforms=None, **kwargs):
    """DDM with trial-to-trial variability: drift v + sv N(0,1), start z-shift 2 (u - 1/2) sz,
    non-decision time t + 2 (u - 1/2) st (ddm_models.pyx:2177-2498). Replay rows:
    replay_normals[i][0] is the drift draw, the walk reads [i][1:]; replay_uniforms is
    [n_total][3] = (u_t, u_z, u_smooth)."""
    return _diffusion("full_ddm", L.MODEL_FULL_DDM, v, a, z, t, deadline, s, max_t, delta_t, n_samples, n_trials,
                      boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads,
                      replay_normals, replay_uniforms, {"sz": sz, "sv": sv, "st": st},
                      {"v": v, "a": a, "z": z, "t": t, "sz": sz, "sv": sv, "st": st, "deadline": deadline, "s": s})


def ornstein_uhlenbeck(v, z, g, t, deadline, s, delta_t=0.001, max_t=20, n_samples=20000, n_trials=1,
                       boundary_fun=None, boundary_params={}, random_state=None, return_option="full",
                       smooth_unif=False, n_threads=1, replay_normals=None, replay_uniforms=None, **kwargs):
    """Ornstein–Uhlenbeck diffusion between ±b(t): step ((v - g y) dt) + sqrt(dt) s N(0,1)
    (ornstein_models.pyx:47-324). The boundary level comes from boundary_params (e.g. "a")."""
    return _diffusion("ornstein_uhlenbeck", L.MODEL_ORNSTEIN, v, None, z, t, deadline, s, max_t, delta_t, n_samples,
                      n_trials, boundary_fun, boundary_params, random_state, return_option, smooth_unif,
                      n_threads, replay_normals, replay_uniforms, {"g": g},
                      {"v": v, "z": z, "g": g, "t": t, "deadline": deadline, "s": s})


def levy_flexbound(v, z, alpha, t, deadline, s, delta_t=0.001, max_t=20, n_samples=20000, n_trials=1,
                   boundary_fun=None, boundary_params={}, random_state=None, return_option="full",
                   smooth_unif=False, n_threads=1, replay_normals=None, replay_uniforms=None, **kwargs):
    """Levy flight between ±b(t): step (v dt) + s dt^(1/alpha) X with X symmetric alpha-stable
    (levy_models.pyx:47-348). The boundary level comes from boundary_params (e.g. "a").
    Replay rows hold the alpha-stable variates themselves."""
    out = _diffusion("levy_flexbound", L.MODEL_LEVY, v, None, z, t, deadline, s, max_t, delta_t, n_samples,
                     n_trials, boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads,
                     replay_normals, replay_uniforms, {"alpha": alpha},
                     {"v": v, "z": z, "t": t, "alpha": alpha, "s": s})
    return out


def racing_diffusion_model(v, b, A, t, s, deadline, delta_t=0.001, max_t=20, n_samples=2000, n_trials=1,
                           random_state=None, return_option="full", smooth_unif=False, n_threads=1,
                           replay_normals=None, replay_uniforms=None, **kwargs):
    """Racing diffusion (Tillman et al. 2020): N Wiener accumulators without a floor, started
    at U(0, A), first to reach b wins (race_models.pyx:925-1300). rt = steps completed before
    the winning step + t (+ smooth_u); no winner by the deadline -> (-999, -1). Replay:
    replay_uniforms[i] = (start u_0 .. u_{N-1}, u_smooth)."""
    _check_threads(n_threads)
    if return_option not in ("full", "minimal"):
        raise ValueError('return_option must be either "full" or "minimal"')
    v = np.asarray(v, np.float32)
    if v.ndim != 2:
        raise ValueError("v must have shape (n_trials, n_particles)")
    n_particles = v.shape[1]
    if n_particles > L.MAX_PARTICLES:
        raise ValueError(f"racing_diffusion_model parallel execution requires n_particles <= {L.MAX_PARTICLES}, "
                         f"got {n_particles}. Use n_threads=1 for larger particle counts.")
    n_trials = int(n_trials)
    col = lambda x: np.asarray(x, np.float32).reshape(n_trials, -1)[:, :1]  # noqa: E731
    max_t = float(np.float32(max_t))
    delta_t = float(np.float32(delta_t))
    rec = accumulator_records(v, col(A), col(t), col(s), deadline, max_t, delta_t, L.BOUNDARY_CONSTANT,
                              {"a": col(b)})
    rts, choices, job, traj = _run(L.MODEL_RDM, L.BOUNDARY_CONSTANT, n_particles, rec, int(n_samples), n_trials,
                                   max_t, delta_t, _seed(random_state), smooth_unif, constant, {"a": col(b)},
                                   replay_normals, replay_uniforms, trajectory=return_option == "full")
    minimal = _minimal_meta("rdm_simulator", list(np.arange(0, n_particles, 1)), int(n_samples), n_trials)
    if return_option == "minimal":
        return {"rts": rts, "choices": choices, "metadata": minimal}
    full = _full_meta(minimal, {"v": v, "b": b, "A": A, "t": t, "deadline": deadline, "s": s},
                      {"delta_t": delta_t, "max_t": max_t, "n_threads": n_threads}, traj=traj,
                      extra={f"v{i}": v[:, i] for i in range(n_particles)})
    return {"rts": rts, "choices": choices, "metadata": full}


def _two_stage(name, model, vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, delta_t, max_t, n_samples, n_trials,
               boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads, replay_normals,
               replay_uniforms):
    _check_threads(n_threads)
    if return_option not in ("full", "minimal"):
        raise ValueError('return_option must be either "full" or "minimal"')
    boundary_fun = boundary_fun or constant
    boundary_params = dict(boundary_params or {})
    if "a" not in boundary_params:
        raise ValueError("boundary_params must provide 'a'")
    code = kernel_boundary_code(boundary_fun)
    max_t = float(np.float32(max_t))
    delta_t = float(np.float32(delta_t))
    n_samples, n_trials = int(n_samples), int(n_trials)
    rec = two_stage_records(vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, max_t, delta_t, code, boundary_params)
    table = boundary_table(boundary_fun, boundary_params, n_trials, max_t, delta_t) if code == L.BOUNDARY_TABLE else None
    job = make_job(model, code, 1, rec, uniform_offsets(n_trials, n_samples), delta_t=delta_t, max_t=max_t,
                   seed=_seed(random_state), smooth_unif=smooth_unif, out_layout=1, choice_bytes=4, table=table,
                   replay_normals=replay_normals, replay_uniforms=replay_uniforms)
    job.launch()
    total = n_samples * n_trials
    shape = (n_samples, n_trials, 1)
    planes = job.rts.cpu().numpy()
    rts = planes[:total].reshape(shape)
    choices = job.choices[:total].cpu().numpy().astype(np.intc).reshape(shape)
    extra = {}
    if model == L.MODEL_PAR2:  # parallel_models.pyx:570: rts_low / rts_high beside rts
        cap = planes.size // 3
        extra = {"rts_low": planes[2 * cap:2 * cap + total].reshape(shape),
                 "rts_high": planes[cap:cap + total].reshape(shape)}
    minimal = _minimal_meta(name, [0, 1, 2, 3], n_samples, n_trials, boundary_fun.__name__)
    if return_option == "minimal":
        return {"rts": rts, "choices": choices, "metadata": minimal, **extra}
    # the reference's two-stage trajectories hold the high walk and both low walks (columns 0-2,
    # sequential_models.pyx:457-579); the kernel walks only the low walk that the high walk
    # selects, so these stay at the -999 fill
    traj = np.full((time_grid(max_t, delta_t).shape[0], 3), -999.0, dtype=np.float32)
    full = _full_meta(
        minimal,
        {"vh": vh, "vl1": vl1, "vl2": vl2, "zh": zh, "zl1": zl1, "zl2": zl2, "t": t, "deadline": deadline, "s": s},
        {"delta_t": delta_t, "max_t": max_t, "n_threads": n_threads}, boundary_fun=boundary_fun,
        boundary=_boundary_trial0(boundary_fun, boundary_params, max_t, delta_t) if n_trials > 0 else None,
        traj=traj, boundary_params=boundary_params,
    )
    return {"rts": rts, "choices": choices, "metadata": full, **extra}


def ddm_flexbound_par2(vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, delta_t=0.001, max_t=20, n_samples=20000,
                       n_trials=1, print_info=True, boundary_fun=None, boundary_params={}, random_state=None,
                       return_option="full", smooth_unif=False, n_threads=1, replay_normals=None,
                       replay_uniforms=None, **kwargs):
    """Hierarchical two-choice DDM, parallel form (parallel_models.pyx:57-603): a high walk
    (vh, zh) and, from t = 0, the low walk its outcome selects (vl1, zl1 or vl2, zl2);
    choices 0..3, rt = max(t_h, t_l) + t (+ smooth_u); returns rts_high / rts_low too.
    Replay: rows of normals hold the high walk's draws then the low walk's; replay_uniforms
    is [n_total][4] = (u_high, unused, u_smooth, u_low)."""
    return _two_stage("ddm_flexbound_par2", L.MODEL_PAR2, vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, delta_t, max_t,
                      n_samples, n_trials, boundary_fun, boundary_params, random_state, return_option, smooth_unif,
                      n_threads, replay_normals, replay_uniforms)


def ddm_flexbound_seq2(vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, delta_t=0.001, max_t=20, n_samples=20000,
                       n_trials=1, print_info=True, boundary_fun=None, boundary_params={}, random_state=None,
                       return_option="full", smooth_unif=False, n_threads=1, replay_normals=None,
                       replay_uniforms=None, **kwargs):
    """Hierarchical two-choice DDM, sequential form (sequential_models.pyx:55-660): the low
    walk continues from where and when the high walk ended; past max_t, or with a low start
    outside the bounds, the low choice comes from its prior zl. Replay: as par2, with
    replay_uniforms[i][1] the prior-choice draw."""
    return _two_stage("ddm_flexbound_seq2", L.MODEL_SEQ2, vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, delta_t, max_t,
                      n_samples, n_trials, boundary_fun, boundary_params, random_state, return_option, smooth_unif,
                      n_threads, replay_normals, replay_uniforms)


def _accumulator(name, model, v, z, t, s, deadline, delta_t, max_t, n_samples, n_trials, boundary_fun,
                 boundary_params, random_state, return_option, smooth_unif, n_threads, g=None, b=None,
                 replay_normals=None, replay_uniforms=None):
    _check_threads(n_threads)
    if return_option not in ("full", "minimal"):
        raise ValueError('return_option must be either "full" or "minimal"')
    v = np.asarray(v, np.float32)
    n_particles = v.shape[1]
    if n_particles > L.MAX_PARTICLES:
        raise ValueError(
            f"{name} parallel execution requires n_particles <= {L.MAX_PARTICLES}, got {n_particles}."
        )
    boundary_fun = boundary_fun or constant
    boundary_params = dict(boundary_params or {})
    code = kernel_boundary_code(boundary_fun)
    max_t = float(np.float32(max_t))
    delta_t = float(np.float32(delta_t))
    rec = accumulator_records(v, z, t, s, deadline, max_t, delta_t, code, boundary_params, g=g, b=b)
    rts, choices, job, traj = _run(model, code, n_particles, rec, int(n_samples), int(n_trials), max_t, delta_t,
                                   _seed(random_state), smooth_unif, boundary_fun, boundary_params, replay_normals,
                                   replay_uniforms, trajectory=return_option == "full")
    minimal = _minimal_meta(name, [-1, 1], int(n_samples), int(n_trials), boundary_fun.__name__)
    if return_option == "minimal":
        return {"rts": rts, "choices": choices, "metadata": minimal}
    minimal["possible_choices"] = list(np.arange(0, n_particles, 1))
    params = {"v": v, "z": z, "t": t, "deadline": deadline, "s": s}
    if g is not None:
        params = {"v": v, "z": z, "g": g, "b": b, "t": t, "deadline": deadline, "s": s}
    extra = {f"v{i}": v[:, i] for i in range(n_particles)}
    zz = np.asarray(z, np.float32)
    extra.update({f"z{i}": zz[:, i] for i in range(zz.shape[1])})
    full = _full_meta(
        minimal, params, {"delta_t": delta_t, "max_t": max_t, "n_threads": n_threads},
        boundary_fun=boundary_fun,
        boundary=_boundary_trial0(boundary_fun, boundary_params, max_t, delta_t) if n_trials > 0 else None,
        traj=traj, boundary_params=boundary_params, extra=extra,
    )
    return {"rts": rts, "choices": choices, "metadata": full}


def race_model(v, z, t, s, deadline, delta_t=0.001, max_t=20, n_samples=2000, n_trials=1, boundary_fun=None,
               boundary_params={}, random_state=None, return_option="full", smooth_unif=False, n_threads=1,
               replay_normals=None, replay_uniforms=None, **kwargs):
    """N-accumulator race, particles clamped at 0 (race_models.pyx:141-506)."""
    return _accumulator("race_model", L.MODEL_RACE, v, z, t, s, deadline, delta_t, max_t, n_samples, n_trials,
                        boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads,
                        replay_normals=replay_normals, replay_uniforms=replay_uniforms)


def lca(v, z, g, b, t, s, deadline, delta_t=0.001, max_t=20, n_samples=2000, n_trials=1, boundary_fun=None,
        boundary_params={}, random_state=None, return_option="full", smooth_unif=False, n_threads=1,
        replay_normals=None, replay_uniforms=None, **kwargs):
    """Leaky competing accumulator (race_models.pyx:509-920)."""
    return _accumulator("lca", L.MODEL_LCA, v, z, t, s, deadline, delta_t, max_t, n_samples, n_trials,
                        boundary_fun, boundary_params, random_state, return_option, smooth_unif, n_threads, g=g,
                        b=b, replay_normals=replay_normals, replay_uniforms=replay_uniforms)


def _lba(name, v, a, z, deadline, sd, t, nact, n_samples, n_trials, max_t, tan_theta, random_state, replay,
         theta=None):
    import ctypes as C  # noqa: PLC0415

    from . import _device as D  # noqa: PLC0415

    torch = D.require_gpu()
    v = np.asarray(v, np.float32).reshape(n_trials, -1)
    rec = np.zeros((n_trials, L.SET_STRIDE), np.float32)
    rec[:, L.P_A] = np.asarray(a, np.float32).reshape(n_trials, -1)[:, 0]
    rec[:, L.P_Z] = np.asarray(z, np.float32).reshape(n_trials, -1)[:, 0]
    rec[:, L.P_T0] = np.asarray(t, np.float32).reshape(-1)
    rec[:, L.P_DEADLINE] = np.asarray(deadline, np.float32).reshape(-1)
    rec[:, L.P_B1] = tan_theta
    rec[:, L.P_VJ : L.P_VJ + nact] = v[:, :nact]
    rec[:, L.P_SSTJ : L.P_SSTJ + nact] = np.broadcast_to(np.asarray(sd, np.float32).reshape(n_trials, -1), (n_trials, nact))
    offs = uniform_offsets(n_trials, n_samples)
    total = int(offs[-1])
    d_rec = D.to_dev(rec)
    d_off = D.to_dev(offs)
    d_rep = D.to_dev(np.ascontiguousarray(replay, np.float32)) if replay is not None else None
    rts = D.empty((max(total, 1),), torch.float32)
    ch = D.empty((max(total, 1),), torch.int32)
    rc = L.load().ssm_lba_simulate(nact, n_trials, total, 1, 4, _seed(random_state), D.ptr(d_rec), D.ptr(d_off),
                                   D.ptr(d_rep), D.ptr(rts), D.ptr(ch), D.stream_ptr())
    L.check(rc, "ssm_lba_simulate")
    rts = rts[:total].cpu().numpy().reshape(n_samples, n_trials, 1)
    choices = ch[:total].cpu().numpy().astype(np.intc).reshape(n_samples, n_trials, 1)
    minimal = _minimal_meta(name, list(np.arange(0, nact, 1)), n_samples, n_trials)
    params = {"a": a, "z": z, "deadline": deadline, "sd": sd, "t": t}
    if theta is not None:  # lba_models.pyx:231
        params = {"a": a, "z": z, "theta": theta, "deadline": deadline, "sd": sd, "t": t}
    meta = _full_meta(minimal, params, {"max_t": max_t, "n_threads": 1},
                      extra={f"v_{i}": v[:, i] for i in range(nact)})
    return {"rts": rts, "choices": choices, "metadata": meta}


def lba_vanilla(v, a, z, deadline, sd, t, nact=3, n_samples=2000, n_trials=1, max_t=20, random_state=None,
                replay=None, **kwargs):
    """Linear ballistic accumulator, closed form (lba_models.pyx:36-139)."""
    return _lba("lba_vanilla", v, a, z, deadline, sd, t, int(nact), int(n_samples), int(n_trials), max_t,
                np.float32(0.0), random_state, replay)


def lba_angle(v, a, z, theta, deadline, sd, t, nact=3, n_samples=2000, n_trials=1, max_t=20, random_state=None,
              replay=None, **kwargs):
    """LBA with an angle-collapsing threshold, closed form (lba_models.pyx:144-250)."""
    th = np.asarray(theta, np.float32).reshape(int(n_trials), -1)[:, 0]
    return _lba("lba_angle", v, a, z, deadline, sd, t, int(nact), int(n_samples), int(n_trials), max_t,
                np.tan(th).astype(np.float32), random_state, replay, theta=theta)
