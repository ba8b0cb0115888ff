"""Host side of the Euler–Maruyama kernels: per-set parameter records and launches.

The per-set record (float32[64]) carries exactly the float32 quantities the
reference computes once per trial before its inner loop:
  ddm_flexbound  src/cssm/ddm_models.pyx:696-717  (boundary, deadline_tmp, sqrt_st)
  race_model     src/cssm/race_models.pyx:378-414
  lca            src/cssm/race_models.pyx:789-823
so the kernel repeats the same float32 operations per step.
"""

from __future__ import annotations

import contextlib
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _device as D
from . import _lib as L
from .boundary_functions import time_grid

F32 = np.float32


def sqrt_dt32(delta_t: float) -> np.float32:
    """float32(sqrt(double(float32 delta_t))) as _utils.pyx:652 / race_models.pyx:379."""
    return F32(np.sqrt(np.float64(F32(delta_t))))


def deadline_tmp32(max_t: float, deadline: np.ndarray, t: np.ndarray) -> np.ndarray:
    """compute_deadline_tmp (_utils.pyx:712-718): min(max_t, deadline - t) in float32."""
    d = np.subtract(np.asarray(deadline, F32), np.asarray(t, F32), dtype=F32)
    return np.minimum(d, F32(max_t), out=d)


def _slope(th: np.ndarray) -> np.ndarray:
    s = np.sin(th)
    np.negative(s, out=s)
    np.divide(s, np.cos(th), out=s)
    return s


def angle_slope32(theta: np.ndarray) -> np.ndarray:
    """-sin(theta)/cos(theta) in float32, the slope boundary_functions.angle forms from each
    trial's scalar theta. NumPy's float32 sin/cos run one SIMD loop with masked tails for
    any length, so the array call gives the per-scalar values (tests/test_host.py checks it).
    Long inputs made of runs of one value (parameters broadcast over trials) are evaluated
    once per run."""
    th = np.asarray(theta, F32).reshape(-1)
    if th.size >= 65536:
        starts = np.flatnonzero(th[1:] != th[:-1]) + 1
        if starts.size < th.size // 8:
            starts = np.concatenate([[0], starts])
            runs = np.diff(np.append(starts, th.size))
            # runs of one length (a parameter broadcast over a trial axis) repeat faster by a count
            return np.repeat(_slope(th[starts]), runs[0] if np.all(runs == runs[0]) else runs)
    return _slope(th)


def boundary_table(fun, params: dict, n_sets: int, max_t: float, delta_t: float) -> np.ndarray:
    """Per-set boundary over t_s, as compute_boundary (_utils.pyx:655-666) builds it per trial."""
    t_s = time_grid(max_t, delta_t)
    out = np.empty((n_sets, t_s.shape[0]), dtype=F32)
    for k in range(n_sets):
        kw = {key: np.asarray(val).reshape(n_sets, -1)[k, 0] if np.ndim(val) > 0 else val for key, val in params.items()}
        out[k, :] = np.asarray(fun(t=t_s, **kw)).astype(F32)
    return out


def new_records(n_sets: int) -> np.ndarray:
    return np.zeros((n_sets, L.SET_STRIDE), dtype=F32)


def first_column(val, n: int) -> np.ndarray:
    """float32[n]: a per-set value given as [n] or [n, 1] (the first column), or a scalar for
    n == 1; empty when n == 0."""
    arr = np.asarray(val, F32)
    return arr.reshape(n, -1)[:, 0] if n else np.zeros(0, F32)


def boundary_columns(code: int, params: dict, n: int) -> dict:
    """{slot: float32 column} of the in-register boundary kinds' parameters."""

    def col(name):
        return first_column(params[name], n)

    if code == L.BOUNDARY_ANGLE:
        return {L.P_B1: angle_slope32(col("theta"))}
    if code == L.BOUNDARY_WEIBULL:
        return {L.P_B1: col("alpha"), L.P_B2: col("beta")}
    return {}


def fill_boundary(rec: np.ndarray, code: int, params: dict) -> None:
    """Write boundary parameters of the in-register boundary kinds into the records."""
    for slot, col in boundary_columns(code, params, rec.shape[0]).items():
        rec[:, slot] = col


def levy_scale32(s, alpha, delta_t) -> np.ndarray:
    """s * pow(delta_t, 1.0 / alpha) (levy_models.pyx:166): float s times a double pow,
    rounded to float once; the scale of the alpha-stable increment over one step."""
    s64 = np.asarray(s, F32).astype(np.float64)
    a64 = np.asarray(alpha, F32).astype(np.float64)
    return (s64 * np.power(np.float64(F32(delta_t)), 1.0 / a64)).astype(F32)


def ddm_columns(v, a, z, t, deadline, s, max_t, delta_t, code, bparams, sz=None, sv=None, st=None, g=None,
                alpha=None) -> tuple[int, dict]:
    """(n, {slot: float32 column}) of the single-particle diffusions' records: every slot they
    use, each formed once per set as the reference forms it per trial (see ddm_records)."""
    v, a, z, t, deadline, s = (np.asarray(x, F32).reshape(-1) for x in (v, a, z, t, deadline, s))
    n = v.shape[0]
    cols = {}
    for slot, val in ((L.P_SZ, sz), (L.P_SV, sv), (L.P_ST, st), (L.P_G, g)):
        if val is not None:
            cols[slot] = first_column(val, n)
    cols[L.P_V] = v
    cols[L.P_Z] = z
    cols[L.P_T0] = t
    cols[L.P_DEADLINE] = deadline
    cols[L.P_DEADLINE_TMP] = deadline_tmp32(max_t, deadline, t)
    cols[L.P_SQRT_ST] = np.multiply(s, sqrt_dt32(delta_t), dtype=F32)
    if alpha is not None:
        al = first_column(alpha, n)
        if np.any(~(al > 0)) or np.any(al > 2):
            raise ValueError("alpha must be in (0, 2]")
        cols[L.P_ALPHA] = al
        cols[L.P_SQRT_ST] = levy_scale32(s, al, delta_t)
    cols[L.P_A] = first_column(bparams.get("a", a), n)
    cols.update(boundary_columns(code, bparams, n))
    return n, cols


def ddm_records(v, a, z, t, deadline, s, max_t, delta_t, code, bparams, sz=None, sv=None, st=None,
                g=None, alpha=None, stride: int = L.SET_STRIDE) -> np.ndarray:
    """Records of the single-particle diffusions: DDM; with sz/sv/st the full DDM
    (ddm_models.pyx:2177-2360); with g the Ornstein–Uhlenbeck leak (ornstein_models.pyx:47-200);
    with alpha the Levy flight (levy_models.pyx:47-220), whose noise scale s dt^(1/alpha)
    replaces sqrt(dt) s. stride = L.SET_STRIDE_SINGLE gives the compact 16-float records these
    models accept (a quarter of the bytes to build and upload)."""
    if stride not in (L.SET_STRIDE, L.SET_STRIDE_SINGLE):
        raise ValueError(f"stride must be {L.SET_STRIDE} or {L.SET_STRIDE_SINGLE}")
    n, cols = ddm_columns(v, a, z, t, deadline, s, max_t, delta_t, code, bparams, sz=sz, sv=sv, st=st, g=g,
                          alpha=alpha)
    return records_from_columns(n, cols, stride)


def records_from_columns(n: int, cols: dict, stride: int) -> np.ndarray:
    """The host layout float32[n, stride] of single-particle {slot: column} records."""
    # the single-particle slots all lie in the first 16 floats: each slot is written as one
    # contiguous row of a slot-major (16, n) array (rec is its transposed view), then the
    # records are laid out by one transposing copy (and, for 64-float records, one more)
    rec = np.zeros((16, n), F32).T
    for slot, col in cols.items():
        rec[:, slot] = col
    if stride == L.SET_STRIDE_SINGLE:
        return np.ascontiguousarray(rec)
    full = new_records(n)
    full[:, :16] = rec
    return full


def records_on_device(n: int, cols: dict, stride: int):
    """The set records as a device tensor float32[n, stride], laid out on the device from the
    columns (ssm_em_records): only the used slots cross to the device, a column at a time,
    and no host transpose is made. Bit-identical to the host layout of the same columns."""
    t = D.require_gpu()
    slots = sorted(cols)
    out = D.empty((max(n, 1), stride), t.float32)
    if n == 0:
        return out[:0]
    dev_cols = D.empty((len(slots), n), t.float32)
    for j, slot in enumerate(slots):
        dev_cols[j].copy_(t.from_numpy(np.ascontiguousarray(cols[slot], F32)))
    arr = (C.c_int32 * len(slots))(*slots)
    rc = L.load().ssm_em_records(n, stride, len(slots), arr, D.ptr(dev_cols), D.ptr(out), D.stream_ptr())
    L.check(rc, "ssm_em_records")
    return out


# below this many sets the records are laid out on the host and uploaded in one copy (a
# copy per slot costs more than the transpose); from it on, on the device (records_on_device)
DEVICE_LAYOUT_MIN_SETS = 4096


def ddm_records_device(*args, stride: int = L.SET_STRIDE_SINGLE, **kwargs):
    """ddm_records as a device tensor float32[n, stride]; same arguments. Bit-identical
    whichever side lays them out."""
    if stride not in (L.SET_STRIDE, L.SET_STRIDE_SINGLE):
        raise ValueError(f"stride must be {L.SET_STRIDE} or {L.SET_STRIDE_SINGLE}")
    n, cols = ddm_columns(*args, **kwargs)
    if n < DEVICE_LAYOUT_MIN_SETS:
        return D.to_dev(records_from_columns(n, cols, stride), F32)
    return records_on_device(n, cols, stride)


def two_stage_records(vh, vl1, vl2, zh, zl1, zl2, t, deadline, s, max_t, delta_t, code, bparams) -> np.ndarray:
    """Records of the hierarchical two-choice DDMs (ddm_flexbound_par2 / _seq2): the high walk
    in the DDM slots (v = vh, z = zh), the two low walks' drifts and starts in the first two
    accumulator slots, and max_t (seq2 stops at the high walk when t_h >= max_t)."""
    vh = np.asarray(vh, F32).reshape(-1)
    n = vh.shape[0]
    rec = ddm_records(vh, bparams.get("a", np.ones(n, F32)), zh, t, deadline, s, max_t, delta_t, code, bparams)
    rec[:, L.P_VJ + 0] = np.asarray(vl1, F32).reshape(-1)
    rec[:, L.P_VJ + 1] = np.asarray(vl2, F32).reshape(-1)
    rec[:, L.P_ZJ + 0] = np.asarray(zl1, F32).reshape(-1)
    rec[:, L.P_ZJ + 1] = np.asarray(zl2, F32).reshape(-1)
    rec[:, L.P_MAX_T] = F32(max_t)
    return rec


def accumulator_records(v, z, t, s, deadline, max_t, delta_t, code, bparams, g=None, b=None) -> np.ndarray:
    v = np.asarray(v, F32)
    z = np.asarray(z, F32)
    s = np.asarray(s, F32)
    n, p = v.shape
    if p > L.MAX_PARTICLES:
        raise ValueError(f"n_particles must be <= {L.MAX_PARTICLES}, got {p}")
    t = np.asarray(t, F32).reshape(n, -1)[:, 0]
    deadline = np.asarray(deadline, F32).reshape(-1)
    rec = new_records(n)
    rec[:, L.P_T0] = t
    rec[:, L.P_DEADLINE] = deadline
    rec[:, L.P_DEADLINE_TMP] = deadline_tmp32(max_t, deadline, t)
    rec[:, L.P_VJ : L.P_VJ + p] = v
    rec[:, L.P_ZJ : L.P_ZJ + p] = np.broadcast_to(z, (n, p))
    rec[:, L.P_SSTJ : L.P_SSTJ + p] = (sqrt_dt32(delta_t) * np.broadcast_to(s, (n, p))).astype(F32)
    if g is not None:
        rec[:, L.P_G] = np.asarray(g, F32).reshape(n, -1)[:, 0]
    if b is not None:
        rec[:, L.P_BINH] = np.asarray(b, F32).reshape(n, -1)[:, 0]
    if "a" in bparams:
        rec[:, L.P_A] = np.asarray(bparams["a"], F32).reshape(n, -1)[:, 0]
    fill_boundary(rec, code, bparams)
    return rec


@dataclass
class LaunchParams:
    block_size: int = 256
    blocks_per_cu: int = 8
    refill_min: int = 8
    queue_block: int = 64  # indices a wave claims per global atomic; 0 = the library default (64)

    def as_struct(self) -> L.LaunchConfig:
        return L.LaunchConfig(self.block_size, self.blocks_per_cu, self.refill_min, self.queue_block)


@dataclass
class EmJob:
    """Everything one persistent launch needs, already resident on the device."""

    prob: L.EmProblem
    records: object  # torch float32 [n_sets, 64]
    offsets: object  # torch int64 [n_sets + 1]
    table: object | None
    replay_normals: object | None
    replay_uniforms: object | None
    rts: object
    choices: object
    workspace: object  # torch int64 [4]
    set_order: object | None = None  # torch int32 [n_sets]: queue visiting order, or None
    queue_offsets: object | None = None  # torch int64 [n_sets + 1]: sizes' prefix in that order

    def launch_config(self, launch: LaunchParams | None = None) -> L.LaunchConfig:
        lc = (launch or LaunchParams()).as_struct()
        lc.set_order = D.ptr(self.set_order)
        lc.queue_offsets = D.ptr(self.queue_offsets)
        return lc

    def set_queue_order(self, order) -> None:
        """Visit the sets in `order` (a permutation of the set indices; None: index order).
        Results do not depend on it; putting long-running sets first shortens the tail."""
        if order is None:
            self.set_order = self.queue_offsets = None
            return
        order = np.asarray(order, np.int64)
        off = self.offsets.cpu().numpy()
        n = off.shape[0] - 1
        if order.shape != (n,) or not np.array_equal(np.sort(order), np.arange(n)):
            raise ValueError("order must be a permutation of the set indices")
        sizes = (off[1:] - off[:-1])[order]
        self.set_order = D.to_dev(order.astype(np.int32))
        self.queue_offsets = D.to_dev(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64))

    def pointers(self):
        return (
            D.ptr(self.records),
            D.ptr(self.offsets),
            D.ptr(self.table),
            D.ptr(self.replay_normals),
            D.ptr(self.replay_uniforms),
            D.ptr(self.rts),
            D.ptr(self.choices),
            D.ptr(self.workspace),
        )

    def launch(self, launch: LaunchParams | None = None, stream=None) -> None:
        lc = self.launch_config(launch)
        rc = L.load().ssm_em_simulate(C.byref(self.prob), C.byref(lc), *self.pointers(), D.stream_ptr(stream))
        L.check(rc, "ssm_em_simulate")

    def trajectory_async(self, n_rows: int, stream=None):
        """Enqueue the replay of walk (set 0, sample 0) on `stream` (default: the current stream)
        and return its device buffer float32[n_rows, n_particles]: the positions after every
        step, the start in row 0, rows after the walk's end -999 (the reference's
        metadata['trajectory']). The walk is replayed on its own draws in one wave; the job's
        outputs and step counters are not touched, so it may run beside the job's launch."""
        t = D.torch()
        p = int(self.prob.n_particles)
        with t.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            traj = t.full((int(n_rows), p), -999.0, dtype=t.float32, device="cuda")
            ws = D.zeros((L.EM_WORKSPACE_WORDS,), t.int64)
            rc = L.load().ssm_em_trajectory(C.byref(self.prob), D.ptr(self.records), D.ptr(self.offsets),
                                            D.ptr(self.table), D.ptr(self.replay_normals),
                                            D.ptr(self.replay_uniforms), D.ptr(traj), int(n_rows), D.ptr(ws),
                                            D.stream_ptr(stream))
        L.check(rc, "ssm_em_trajectory")
        return traj

    def trajectory(self, n_rows: int, stream=None) -> np.ndarray:
        """trajectory_async, waited for and copied to the host."""
        traj = self.trajectory_async(n_rows, stream)
        if stream is not None:
            stream.synchronize()
        return traj.cpu().numpy()

    def steps(self) -> int:
        """EM steps executed by the last completed launch (device counter)."""
        return int(self.workspace[1].item())

    def lane_slots(self) -> int:
        """Lane-step slots issued in the last launch (64 lanes x the steps of each wave
        iteration); steps() / lane_slots() is the fraction of lanes doing useful steps."""
        return int(self.workspace[2].item())


def expected_steps(records: np.ndarray, kind: int, n_particles: int, max_t: float, delta_t: float) -> np.ndarray:
    """A cheap per-set proxy of the mean walk length (in steps), used only to order the queue:
    the mean exit time of a Wiener process with drift v and unit noise from (-a, a) started at
    -a + 2 z a (constant bound) for the diffusions; a / max_j v_j for the accumulators."""
    rec = np.asarray(records, np.float64)
    a = np.maximum(rec[:, L.P_A], 1e-6)
    if kind in (L.MODEL_DDM, L.MODEL_FULL_DDM, L.MODEL_ORNSTEIN, L.MODEL_LEVY, L.MODEL_PAR2, L.MODEL_SEQ2,
                L.MODEL_DDM_LEGACY):
        v = rec[:, L.P_V]  # two-stage models: the high walk's
        s2 = np.maximum((rec[:, L.P_SQRT_ST] ** 2) / delta_t, 1e-12)  # noise variance s^2
        A_, x0 = 2 * a, 2 * a * np.clip(rec[:, L.P_Z], 0.0, 1.0)
        mu = v / s2
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            e1 = np.expm1(np.clip(-2 * mu * x0, -700, 700))
            e2 = np.expm1(np.clip(-2 * mu * A_, -700, 700))
            t = (A_ * e1 / e2 - x0) / v
        t = np.where(np.abs(v) < 1e-6 * np.sqrt(s2) / A_, x0 * (A_ - x0) / s2, t)
    else:
        vmax = rec[:, L.P_VJ:L.P_VJ + n_particles].max(axis=1)
        t = np.where(vmax > 1e-6, a / np.maximum(vmax, 1e-6), max_t)
    t = np.nan_to_num(t, nan=max_t, posinf=max_t)
    return np.clip(t, 0.0, max_t) / delta_t


def make_job(
    model: int,
    boundary: int,
    n_particles: int,
    records,
    offsets,
    *,
    delta_t: float,
    max_t: float,
    seed: int,
    smooth_unif: bool,
    out_layout: int = 0,
    choice_bytes: int = 4,
    table: np.ndarray | None = None,
    replay_normals: np.ndarray | None = None,
    replay_uniforms: np.ndarray | None = None,
) -> EmJob:
    """One persistent launch's inputs and outputs on the device. records: float32[n_sets, stride]
    on the host, or already on the device (records_on_device). offsets: int64[n_sets + 1] on
    the host, or an int n: every set has n samples, and the offsets are written on the device."""
    t = D.require_gpu()
    on_device = not isinstance(records, np.ndarray) and bool(getattr(records, "is_cuda", False))
    if on_device and (records.dtype != t.float32 or not records.is_contiguous()):
        raise ValueError("device records must be a contiguous float32 tensor")
    if records.ndim != 2 or records.shape[1] not in (L.SET_STRIDE, L.SET_STRIDE_SINGLE):
        raise ValueError(f"records must have {L.SET_STRIDE} (or, single-particle diffusions, "
                         f"{L.SET_STRIDE_SINGLE}) floats per set")
    n_sets = int(records.shape[0])
    uniform = isinstance(offsets, (int, np.integer))
    if uniform:
        per_set = int(offsets)
        if per_set < 0:
            raise ValueError("n_samples per set must be >= 0")
        n_total = n_sets * per_set
    else:
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        if offsets.shape != (n_sets + 1,):
            raise ValueError("offsets must have n_sets + 1 entries")
        if np.any(np.diff(offsets) < 0) or offsets[0] != 0:
            raise ValueError("offsets must start at 0 and be non-decreasing")
        n_total = int(offsets[-1])
    from .boundary_functions import n_steps as _n_steps

    nst = _n_steps(max_t, delta_t)
    stride = 0
    if replay_normals is not None:
        replay_normals = np.ascontiguousarray(replay_normals, dtype=F32)
        if replay_normals.ndim != 2 or replay_normals.shape[0] != n_total:
            raise ValueError("replay_normals must have shape (n_total, stride)")
        stride = int(replay_normals.shape[1])
        if stride < 1:
            raise ValueError("replay stride must be >= 1")
    if replay_uniforms is not None:
        # values per trajectory: full DDM (u_t, u_z, u_smooth), racing diffusion (u_start x P, u_smooth)
        per = {L.MODEL_FULL_DDM: 3, L.MODEL_RDM: n_particles + 1, L.MODEL_PAR2: 4, L.MODEL_SEQ2: 4}.get(model, 1)
        replay_uniforms = np.ascontiguousarray(replay_uniforms, dtype=F32)
        if replay_uniforms.size != n_total * per:
            raise ValueError(f"replay_uniforms must hold {per} value(s) per trajectory ({n_total * per})")
    if stride and model == L.MODEL_RDM and replay_uniforms is None:
        raise ValueError("racing-diffusion replay needs replay_uniforms (start draws and u_smooth)")
    if stride and model in (L.MODEL_PAR2, L.MODEL_SEQ2) and replay_uniforms is None:
        raise ValueError("two-stage replay needs replay_uniforms (four per trajectory)")
    if table is not None:
        table = np.ascontiguousarray(table, dtype=F32)
        if table.shape != (n_sets, nst):
            raise ValueError(f"boundary table must have shape ({n_sets}, {nst})")
    prob = L.EmProblem(
        model=model,
        boundary=boundary,
        n_particles=n_particles,
        n_sets=n_sets,
        n_total=n_total,
        n_steps=nst,
        dt=float(F32(delta_t)),
        smooth_unif=int(bool(smooth_unif)),
        out_layout=out_layout,
        choice_bytes=choice_bytes,
        set_stride=0 if records.shape[1] == L.SET_STRIDE else int(records.shape[1]),
        seed=int(seed) & 0xFFFFFFFFFFFFFFFF,
        replay_stride=stride,
    )
    cdt = t.int8 if choice_bytes == 1 else t.int32
    return EmJob(
        prob=prob,
        records=records if on_device else D.to_dev(records, F32),
        offsets=uniform_offsets_device(n_sets, per_set) if uniform else D.to_dev(offsets, np.int64),
        table=D.to_dev(table) if table is not None else None,
        replay_normals=D.to_dev(replay_normals) if replay_normals is not None else None,
        replay_uniforms=D.to_dev(np.ascontiguousarray(replay_uniforms, F32)) if replay_uniforms is not None else None,
        # par2 also returns the high and the low walk's rts: (rt, rt_high, rt_low) planes
        rts=D.empty(((3 if model == L.MODEL_PAR2 else 1) * max(n_total, 1),), t.float32),
        choices=D.empty((max(n_total, 1),), cdt),
        workspace=D.zeros((L.EM_WORKSPACE_WORDS,), t.int64),
    )


def uniform_offsets_device(n_sets: int, n_samples: int):
    """int64[n_sets + 1] device tensor k * n_samples, written on the device (ssm_uniform_offsets)."""
    t = D.require_gpu()
    out = D.empty((n_sets + 1,), t.int64)
    L.check(L.load().ssm_uniform_offsets(int(n_sets), int(n_samples), D.ptr(out), D.stream_ptr()),
            "ssm_uniform_offsets")
    return out


def uniform_offsets(n_sets: int, n_samples: int) -> np.ndarray:
    return (np.arange(n_sets + 1, dtype=np.int64) * int(n_samples)).astype(np.int64)


def boundary_table_for(fun, bparams, n_sets, max_t, delta_t):
    return boundary_table(fun, bparams, n_sets, max_t, delta_t)


def time_grid_for(max_t, delta_t):
    return time_grid(max_t, delta_t)
