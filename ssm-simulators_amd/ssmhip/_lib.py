"""ctypes binding of libssmhip.so (the C ABI declared in include/ssmhip.h).

The library is built in-tree by ``ssm-simulators_amd/csrc/build.py``. There is
no CPU fallback: a missing library or a missing GPU raises immediately.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

NATIVE_DIR = Path(__file__).resolve().parent / "_native"
LIB_PATH = NATIVE_DIR / "libssmhip.so"

# status codes (ssmhip.h)
SSM_OK = 0
SSM_EINVAL = -1
SSM_EHIP = -2
SSM_ENOMEM = -3
SSM_EUNSUPPORTED = -4

MODEL_DDM, MODEL_RACE, MODEL_LCA, MODEL_FULL_DDM, MODEL_ORNSTEIN, MODEL_LEVY, MODEL_RDM = 0, 1, 2, 3, 4, 5, 6
MODEL_PAR2, MODEL_SEQ2, MODEL_DDM_LEGACY = 7, 8, 9
BOUNDARY_CONSTANT, BOUNDARY_ANGLE, BOUNDARY_WEIBULL, BOUNDARY_TABLE = 0, 1, 2, 3
MAX_PARTICLES = 16
SET_STRIDE = 64
SET_STRIDE_SINGLE = 16  # compact records of the single-particle diffusions (slots 0..15)
P_T0, P_DEADLINE, P_DEADLINE_TMP, P_A, P_B1, P_B2, P_Z, P_V, P_SQRT_ST, P_G, P_BINH, P_SZ, P_SV, P_ST, P_ALPHA, P_MAX_T = range(16)
P_VJ, P_ZJ, P_SSTJ = 16, 32, 48
EM_WORKSPACE_WORDS = 4
KDE_CHUNK = 8192
KDE_EVAL_SCRATCH_FLOATS = 16 * 1024 * 1024  # SSM_KDE_EVAL_SCRATCH_FLOATS

class EmProblem(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("boundary", C.c_int32),
        ("n_particles", C.c_int32),
        ("n_sets", C.c_int32),
        ("n_total", C.c_int64),
        ("n_steps", C.c_int32),
        ("dt", C.c_float),
        ("smooth_unif", C.c_int32),
        ("out_layout", C.c_int32),
        ("choice_bytes", C.c_int32),
        ("set_stride", C.c_int32),
        ("seed", C.c_uint64),
        ("replay_stride", C.c_int64),
    ]


class LaunchConfig(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("refill_min", C.c_int32),
        ("queue_block", C.c_int32),
        ("set_order", C.c_void_p),
        ("queue_offsets", C.c_void_p),
    ]


class SsmError(RuntimeError):
    """A HIP/runtime failure reported by libssmhip."""


_lock = threading.Lock()
_lib: C.CDLL | None = None

_VP = C.c_void_p
_SIGS = {
    "ssm_abi_version": (C.c_int32, []),
    "ssm_last_error": (C.c_char_p, []),
    "ssm_device_cu_count": (C.c_int32, []),
    "ssm_em_simulate": (
        C.c_int32,
        [C.POINTER(EmProblem), C.POINTER(LaunchConfig), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "ssm_em_trajectory": (C.c_int32, [C.POINTER(EmProblem), _VP, _VP, _VP, _VP, _VP, _VP, C.c_int32, _VP, _VP]),
    "ssm_em_graph_create": (
        C.c_int32,
        [C.POINTER(EmProblem), C.POINTER(LaunchConfig), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(_VP)],
    ),
    "ssm_graph_launch": (C.c_int32, [_VP, _VP]),
    "ssm_graph_destroy": (C.c_int32, [_VP]),
    "ssm_lba_simulate": (
        C.c_int32,
        [C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_uint64, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "ssm_kde_fit": (C.c_int32, [C.c_int32, C.c_int32, _VP, _VP, _VP, C.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "ssm_kde_sample": (
        C.c_int32,
        [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_uint64, _VP, _VP, _VP, _VP, _VP, _VP,
         _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "ssm_kde_eval": (
        C.c_int32,
        [C.c_int32, C.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_float, C.c_float,
         C.c_int64, _VP, _VP, _VP],
    ),
    "ssm_kde_eval_noalloc": (
        C.c_int32,
        [C.c_int32, C.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_float, C.c_float, C.c_float,
         C.c_int64, C.c_int64, _VP, _VP],
    ),
    "ssm_lan_assemble": (
        C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_uint64, _VP,
                    _VP, _VP, _VP, _VP, _VP, _VP]
    ),
    "ssm_copy_rows": (C.c_int32, [C.c_int64, C.c_int64, _VP, _VP, _VP, _VP, _VP]),
    "ssm_em_records": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _VP, _VP, _VP, _VP]),
    "ssm_uniform_offsets": (C.c_int32, [C.c_int32, C.c_int64, _VP, _VP]),
    "ssm_mlp_sizes": (C.c_int32, [C.POINTER(C.c_int32)] * 5),
    "ssm_mlp_pack": (C.c_int32, [_VP, _VP, _VP, _VP]),
    "ssm_mlp_forward": (C.c_int32, [_VP, C.c_int64, _VP, _VP, _VP, _VP, _VP]),
    "ssm_mlp_backward": (
        C.c_int32,
        [_VP, _VP, C.c_int64, C.c_int32, C.c_float, C.c_float, _VP, _VP, _VP, _VP, _VP, C.c_int64, _VP, _VP, _VP,
         _VP],
    ),
    "ssm_mlp_wgrad": (C.c_int32, [_VP, _VP, C.c_int64, C.c_int64, _VP, _VP]),
    "ssm_mlp_adam": (
        C.c_int32, [_VP, _VP, _VP, _VP, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _VP]
    ),
    "ssm_mlp_train_workspace_bytes": (C.c_int64, [C.c_int64]),
    "ssm_mlp_train_steps": (
        C.c_int32,
        [_VP, _VP, C.c_int64, C.c_int32, C.c_int32, _VP, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
         _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int64, _VP, _VP],
    ),
    "ssm_aux_labels": (
        C.c_int32,
        [C.c_int32, C.c_int32, _VP, C.c_int32, _VP, _VP, C.c_int32, _VP, C.c_double, _VP, _VP, _VP, _VP, _VP],
    ),
    "ssm_theta_uniform": (C.c_int32, [C.c_int32, C.c_int32, _VP, _VP, C.c_uint64, _VP, _VP]),
    "ssm_filter_workspace_bytes": (C.c_int32, [C.c_int32, C.c_int64, C.POINTER(C.c_size_t)]),
    "ssm_filter_stats": (
        C.c_int32, [C.c_int32, C.c_int32, _VP, _VP, _VP, C.c_int32, _VP, C.c_int64, _VP, C.c_size_t, _VP, _VP]
    ),
    "ssm_mfma_probe": (C.c_int32, [_VP, _VP, _VP, _VP]),
    "ssm_mfma32_probe": (C.c_int32, [_VP, _VP, _VP, _VP]),
    "ssm_valu_probe": (C.c_int32, [C.c_int32, C.c_int64, C.c_int64, _VP, _VP]),
    "ssm_philox_fill": (
        C.c_int32,
        [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int64, _VP, _VP, _VP],
    ),
}


def lib_path() -> Path:
    return Path(os.environ.get("SSMHIP_LIB", str(LIB_PATH)))


def load(require: bool = True) -> C.CDLL | None:
    """Load libssmhip.so once; raise (require=True) when it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not path.exists():
            if require:
                raise ImportError(
                    f"libssmhip.so not found at {path}; build it with "
                    "`python ssm-simulators_amd/csrc/build.py` (or __graft_entry__.build())"
                )
            return None
        lib = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    lib = load()
    msg = lib.ssm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Turn a libssmhip status into the reference's exception types."""
    if rc == SSM_OK:
        return
    msg = last_error()
    if what:
        msg = f"{what}: {msg}"
    if rc == SSM_EINVAL:
        raise ValueError(msg)
    raise SsmError(msg)


def fn(name: str):
    lib = load()
    f = getattr(lib, name)
    if name not in _SIGS and not getattr(f, "_ssm_typed", False):
        f.restype = C.c_int32
        f._ssm_typed = True
    return f
