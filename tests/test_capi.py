"""The C ABI library loads and exports every entry point include/ssmhip.h declares (CPU, no GPU calls)."""

import ctypes as C
import re
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "ssmhip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ssm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "ssm_em_simulate" in names and "ssm_last_error" in names
    assert len(names) >= 8


def test_library_exports_every_declared_symbol():
    from ssmhip import _lib

    if not _lib.lib_path().exists():
        pytest.fail(f"libssmhip.so is not built at {_lib.lib_path()}; run __graft_entry__.build()")
    lib = C.CDLL(str(_lib.lib_path()))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_abi_version_and_error_string_without_gpu():
    from ssmhip import _lib

    lib = _lib.load()
    assert lib.ssm_abi_version() == 6
    assert isinstance(lib.ssm_last_error(), bytes)


def test_em_rejects_bad_descriptor_before_touching_the_gpu():
    from ssmhip import _lib

    lib = _lib.load()
    prob = _lib.EmProblem(model=99, boundary=0, n_particles=1, n_sets=1, n_total=1, n_steps=10, dt=0.001)
    rc = lib.ssm_em_simulate(C.byref(prob), None, *([None] * 8), None)
    assert rc == _lib.SSM_EINVAL
    assert "unknown model" in _lib.last_error()
    prob = _lib.EmProblem(model=0, boundary=0, n_particles=2, n_sets=1, n_total=1, n_steps=10, dt=0.001)
    assert lib.ssm_em_simulate(C.byref(prob), None, *([None] * 8), None) == _lib.SSM_EINVAL
    assert "n_particles" in _lib.last_error()
    with pytest.raises(ValueError, match="n_particles"):
        _lib.check(_lib.SSM_EINVAL)


def test_kde_eval_noalloc_rejects_sizes_beyond_its_scratch_before_touching_the_gpu():
    from ssmhip import _lib

    lib = _lib.load()
    nch = 13  # ceil(1e5 / 8192) partial sums per point
    rc = lib.ssm_kde_eval_noalloc(1, 2, *([None] * 9), C.c_float(20.0), C.c_float(-66.774), C.c_float(1e-4),
                                  100_000, _lib.KDE_EVAL_SCRATCH_FLOATS // nch + 1, None, None)
    assert rc == _lib.SSM_EINVAL
    assert "no-allocation scratch" in _lib.last_error()
    assert lib.ssm_kde_eval_noalloc(1, 2, *([None] * 9), C.c_float(20.0), C.c_float(-66.774), C.c_float(1e-4),
                                    100_000, -1, None, None) == _lib.SSM_EINVAL


def test_records_and_offsets_reject_bad_arguments_before_touching_the_gpu():
    from ssmhip import _lib

    lib = _lib.load()
    slots = (C.c_int32 * 2)(3, 3)
    assert lib.ssm_em_records(4, 16, 2, slots, None, None, None) == _lib.SSM_EINVAL
    slots = (C.c_int32 * 2)(3, 3)
    x = C.c_float(0.0)
    assert lib.ssm_em_records(4, 16, 2, slots, C.byref(x), C.byref(x), None) == _lib.SSM_EINVAL
    assert "twice" in _lib.last_error()
    slots = (C.c_int32 * 1)(16,)
    assert lib.ssm_em_records(4, 16, 1, slots, C.byref(x), C.byref(x), None) == _lib.SSM_EINVAL
    assert "outside the record" in _lib.last_error()
    assert lib.ssm_em_records(4, 65, 0, None, None, C.byref(x), None) == _lib.SSM_EINVAL
    assert lib.ssm_em_records(0, 16, 0, None, None, None, None) == _lib.SSM_OK
    assert lib.ssm_uniform_offsets(-1, 5, None, None) == _lib.SSM_EINVAL
    assert lib.ssm_uniform_offsets(3, 5, None, None) == _lib.SSM_EINVAL
