"""ssm_kde_eval_noalloc (partial sums in the library's fixed device scratch) gives exactly what
ssm_kde_eval (caller's workspace) gives, on two calls in a row, the second one smaller."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


def _sets(n_sets, n, seed):
    rng = np.random.default_rng(seed)
    rts, chs = [], []
    for k in range(n_sets):
        m = int(rng.integers(n // 4, n + 1))
        ch = np.where(rng.uniform(size=m) < rng.uniform(0.1, 0.9), 1, -1).astype(np.int8)
        rt = (rng.uniform(0.1, 1.0) + rng.lognormal(-0.5, 0.6, size=m)).astype(F32)
        rt[rng.uniform(size=m) < 0.01] = -999.0
        rts.append(rt)
        chs.append(ch)
    return rts, chs


def _both(n_sets, n, n_kde, n_unif, seed):
    import torch

    from ssmhip import _device as D
    from ssmhip import _lib as L
    from ssmhip.kde import LB_DEFAULT, KDEBank

    rts, chs = _sets(n_sets, n, seed)
    offs = np.concatenate([[0], np.cumsum([r.size for r in rts])]).astype(np.int64)
    bank = KDEBank(D.to_dev(np.concatenate(rts)), D.to_dev(np.concatenate(chs)), D.to_dev(offs), [-1, 1])
    pts = bank.sample(n_kde, n_unif, seed=seed)
    ref = bank.evaluate(pts)  # ssm_kde_eval with a workspace from the caller
    n_points = int(pts.rt.shape[0])
    out = D.empty((n_points,), torch.float32)
    out.fill_(np.nan)
    print(f"[case] kde_eval_noalloc n_sets={n_sets} n_points={n_points} max_samples={bank.max_samples}", flush=True)
    rc = L.load().ssm_kde_eval_noalloc(bank.n_sets, bank.C, D.ptr(pts.grp_begin), D.ptr(pts.grp_count),
                                       D.ptr(pts.rt), D.ptr(bank.samples), D.ptr(bank.kde_count),
                                       D.ptr(bank.choice_total), D.ptr(bank.bandwidth), D.ptr(bank.kde_start),
                                       D.ptr(bank.offsets), C.c_float(20.0), C.c_float(LB_DEFAULT),
                                       C.c_float(10e-5), bank.max_samples, n_points, D.ptr(out), D.stream_ptr())
    L.check(rc, "ssm_kde_eval_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_eval_equals_workspace_eval_on_consecutive_calls(torch_gpu):
    # first call: 96 sets of up to 40,000 draws (5 chunks of partial sums per point)
    ref1, got1 = _both(96, 40_000, 800, 200, seed=3)
    assert np.all(np.isfinite(ref1))
    assert np.array_equal(ref1, got1)
    # second call on the same scratch: fewer sets, fewer draws, fewer points
    ref2, got2 = _both(7, 9_000, 300, 50, seed=4)
    assert np.all(np.isfinite(ref2))
    assert np.array_equal(ref2, got2)


def test_noalloc_eval_rejects_points_beyond_the_scratch(torch_gpu):
    from ssmhip import _lib as L

    nch = 13  # ceil(1e5 / 8192)
    too_many = L.KDE_EVAL_SCRATCH_FLOATS // nch + 1
    rc = L.load().ssm_kde_eval_noalloc(1, 2, *([None] * 9), C.c_float(20.0), C.c_float(-66.774), C.c_float(1e-4),
                                       100_000, too_many, None, None)
    assert rc == L.SSM_EINVAL
    assert "no-allocation scratch" in L.last_error()
