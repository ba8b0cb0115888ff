"""ctypes loader of oracle/_lib/libssm_oracle.so (test infrastructure only).

Used by tests (as a fast CPU oracle whose replay mode is pinned bit-exactly
against em_numpy) and by bench.py's cpu_baseline leg (the reported all-core
CPU baseline: MT19937 + Ziggurat, OpenMP dynamic over (set, sample)).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "_lib" / "libssm_oracle.so"
F32 = np.float32

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        P = C.c_void_p
        lib.cpu_em_simulate.restype = C.c_int64
        lib.cpu_em_simulate.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, C.c_int, C.c_float,
                                        C.c_int, C.c_uint64, P, C.c_int64, P, P, P, C.c_int]
        lib.cpu_ziggurat_normals.restype = None
        lib.cpu_ziggurat_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, P]
        lib.cpu_mix_seed.restype = C.c_uint64
        lib.cpu_mix_seed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        lib.cpu_kde_logsum.restype = None
        lib.cpu_kde_logsum.argtypes = [P, C.c_int64, P, C.c_int64, C.c_double, P, C.c_int]
        _lib = lib
    return _lib


def n_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def em_simulate(model: int, v, z, sst, t0, deadline, dl_tmp, table, offsets, *, dt, seed=0, g=None, b=None,
                smooth_unif=False, replay_normals=None, replay_uniforms=None, n_threads=0):
    """Run the C++ restatement. v/z/sst: float32[K, P]; per-set vectors float32[K]; table float32[K, n_steps]."""
    lib = load()
    v = np.ascontiguousarray(v, F32)
    K = v.shape[0]
    P = 1 if v.ndim == 1 else v.shape[1]
    v = v.reshape(K, P)
    z = np.ascontiguousarray(np.broadcast_to(np.asarray(z, F32).reshape(K, -1), (K, P)))
    sst = np.ascontiguousarray(np.broadcast_to(np.asarray(sst, F32).reshape(K, -1), (K, P)))
    t0, deadline, dl_tmp = (np.ascontiguousarray(np.asarray(a, F32).reshape(K)) for a in (t0, deadline, dl_tmp))
    g = np.ascontiguousarray(np.zeros(K, F32) if g is None else np.asarray(g, F32).reshape(K))
    b = np.ascontiguousarray(np.zeros(K, F32) if b is None else np.asarray(b, F32).reshape(K))
    table = np.ascontiguousarray(table, F32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    total = int(offsets[-1])
    rts = np.empty(total, F32)
    ch = np.empty(total, np.int32)
    stride = 0
    if replay_normals is not None:
        replay_normals = np.ascontiguousarray(replay_normals, F32)
        stride = replay_normals.shape[1]
    if replay_uniforms is not None:
        replay_uniforms = np.ascontiguousarray(replay_uniforms, F32)
    steps = lib.cpu_em_simulate(model, P, K, _p(offsets), _p(v), _p(z), _p(sst), _p(t0), _p(deadline), _p(dl_tmp),
                                _p(g), _p(b), _p(table), table.shape[1], F32(dt), int(bool(smooth_unif)), seed,
                                _p(replay_normals), stride, _p(replay_uniforms), _p(rts), _p(ch), int(n_threads))
    return rts, ch, int(steps)


def ziggurat_normals(seed: int, k: int, n: int, count: int) -> np.ndarray:
    out = np.empty(count, F32)
    load().cpu_ziggurat_normals(seed, k, n, count, _p(out))
    return out


def mix_seed(base: int, t1: int, t2: int) -> int:
    return int(load().cpu_mix_seed(base, t1, t2))


def kde_logsum(x: np.ndarray, samples: np.ndarray, h: float, n_threads: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    s = np.ascontiguousarray(samples, np.float64)
    out = np.empty(x.shape[0], np.float64)
    load().cpu_kde_logsum(_p(x), x.shape[0], _p(s), s.shape[0], float(h), _p(out), int(n_threads))
    return out
