"""Auxiliary labels and the parameter sampler on the GPU, against the oracle (exact counts)."""

import numpy as np
import pytest

from oracle import aux_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("choice_bytes", [1, 4])
def test_aux_counts_match_oracle_exactly(torch_gpu, choice_bytes):
    from ssmhip import _device as D
    from ssmhip import _lib as L

    torch = torch_gpu
    rng = np.random.default_rng(1)
    sizes = np.array([0, 1, 7, 5000, 333, 0, 2048])
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs[-1])
    rts = rng.uniform(-1.0, 22.0, n).astype(np.float32)
    edges = np.concatenate([np.arange(129) * 0.15625, np.arange(257) * 0.078125]).astype(np.float32)
    rts[: edges.size] = edges  # values exactly on bin edges, at 0 and at max_t
    rts[rng.random(n) < 0.05] = -999.0
    vals = [0, 1, 2]
    ch = rng.choice([0, 1, 2, 3], n, p=[0.3, 0.3, 0.3, 0.1]).astype(np.int8 if choice_bytes == 1 else np.int32)
    dc = D.to_dev(ch)
    S, C = sizes.size, len(vals)
    out = {k: D.empty(s, torch.int32) for k, s in
           [("c", (S, C)), ("ng", (S,)), ("h1", (S, 128, C)), ("h2", (S, 256, C))]}
    # device buffers held in names: a temporary freed before the launch could be reused
    d_vals, d_rts, d_offs = D.to_dev(np.array(vals, np.int32)), D.to_dev(rts), D.to_dev(offs)
    rc = L.load().ssm_aux_labels(S, C, D.ptr(d_vals), 2, D.ptr(d_rts), D.ptr(dc), choice_bytes, D.ptr(d_offs), 20.0,
                                 D.ptr(out["c"]), D.ptr(out["ng"]), D.ptr(out["h1"]), D.ptr(out["h2"]),
                                 D.stream_ptr())
    L.check(rc)
    got = {k: v.cpu().numpy() for k, v in out.items()}
    for k in range(S):
        lo, hi = offs[k], offs[k + 1]
        if hi == lo:
            assert got["c"][k].sum() == 0 and got["ng"][k] == 0 and got["h1"][k].sum() == 0
            continue
        ref = aux_ref.labels_one_set(rts[lo:hi], ch[lo:hi].astype(np.int32), vals, 20.0)
        m = hi - lo
        assert np.array_equal(got["c"][k] / m, ref["choice_p"][0])
        assert got["ng"][k] / m == ref["nogo_p"][0, 0]
        assert np.array_equal(got["h1"][k], ref["binned_128"][0].astype(np.int64))
        assert np.array_equal(got["h2"][k], ref["binned_256"][0].astype(np.int64))


def test_labeller_output_dict_matches_oracle(torch_gpu):
    from ssmhip.lan import LanKDELabeller, sample_angle_uniform

    theta = sample_angle_uniform(8, seed=3)
    lab = LanKDELabeller("angle", theta, n_samples=20000, launch=None, seed=9)
    rows = lab.run()
    out = lab.output_dict(rows)
    rts = lab.sim.rts.cpu().numpy()
    ch = lab.sim.choices.cpu().numpy().astype(np.int32)
    ref = aux_ref.labels_all_sets(rts, ch, lab.sim.offsets, [-1, 1], 20.0)
    for key, val in ref.items():
        assert out[key].dtype == np.float32 and out[key].shape == val.shape, key
        assert np.array_equal(out[key], val), key
    assert out["lan_data"].shape == (8 * 1000, 7) and out["lan_labels"].shape == (8 * 1000,)
    for key in ("cpn_data", "opn_data", "gonogo_data", "cpn_no_omission_data", "theta"):
        assert np.array_equal(out[key], theta)


def test_theta_sampler_bounds_determinism_and_uniformity(torch_gpu):
    from scipy import stats

    from ssmhip.lan import sample_theta

    lo = [-3.0, 0.3, 0.1, 1e-3, -0.1]
    hi = [3.0, 3.0, 0.9, 2.0, 1.3]
    a = sample_theta(20000, lo, hi, seed=5)
    b = sample_theta(20000, lo, hi, seed=5)
    c = sample_theta(20000, lo, hi, seed=6)
    assert a.dtype == np.float32 and a.shape == (20000, 5)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for j in range(5):
        assert a[:, j].min() >= np.float32(lo[j]) and a[:, j].max() <= np.float32(hi[j])
        assert stats.kstest((a[:, j] - lo[j]) / (hi[j] - lo[j]), "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(a[:, 0], a[:, 1])[0, 1]) < 0.05
    with pytest.raises(ValueError):
        sample_theta(10, [1.0], [0.0])


@pytest.mark.parametrize("choice_bytes", [1, 4])
def test_filter_stats_match_oracle(torch_gpu, choice_bytes):
    """Per-set, per-choice (mode, mean, std, mode share, count) of ragged sets against the
    scipy/NumPy restatement: counts, modes and shares exact (ties: the smaller value), mean and
    std within float32 rounding."""
    import ctypes

    from ssmhip import _device as D
    from ssmhip import _lib as L

    torch = torch_gpu
    rng = np.random.default_rng(5)
    sizes = np.array([0, 1, 4, 5, 3000, 777, 0, 20000])
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs[-1])
    # a coarse grid gives ties (several values sharing the top count), plus continuous values
    rts = np.where(rng.random(n) < 0.5, rng.integers(0, 40, n) * np.float32(0.025),
                   rng.uniform(0.0, 3.0, n)).astype(np.float32)
    rts[rng.random(n) < 0.03] = -999.0
    vals = [-1, 1, 7]  # 7 never occurs: an empty choice in every set
    ch = rng.choice([-1, 1, 3], n, p=[0.45, 0.45, 0.1]).astype(np.int8 if choice_bytes == 1 else np.int32)
    S, C = sizes.size, len(vals)
    d_vals, d_rts, d_ch, d_offs = (D.to_dev(x) for x in (np.array(vals, np.int32), rts, ch, offs))
    nb = ctypes.c_size_t(0)
    L.check(L.load().ssm_filter_workspace_bytes(S, n, ctypes.byref(nb)))
    ws = D.empty((nb.value,), torch.uint8)
    stats = D.empty((S, C, 5), torch.float64)
    rc = L.load().ssm_filter_stats(S, C, D.ptr(d_vals), D.ptr(d_rts), D.ptr(d_ch), choice_bytes, D.ptr(d_offs), n,
                                   D.ptr(ws), nb.value, D.ptr(stats), D.stream_ptr())
    L.check(rc)
    got = stats.cpu().numpy()
    for k in range(S):
        ref = aux_ref.filter_stats_one_set(rts[offs[k]:offs[k + 1]], ch[offs[k]:offs[k + 1]], vals)
        assert np.array_equal(got[k][:, [0, 3, 4]], ref[:, [0, 3, 4]]), k
        assert np.allclose(got[k][:, 1:3], ref[:, 1:3], rtol=2e-6, atol=1e-6), k
    assert np.all(got[:, 2, :] == [-1, -1, 1, 0, 0])


def test_filter_resamples_rejected_sets_until_they_pass(torch_gpu):
    """A filter on the mean rt: rejected parameter sets are re-drawn and re-simulated; every
    returned set passes, and an impossible filter drops all sets."""
    from ssmhip.lan import LanKDELabeller, sample_theta

    lo, hi = [-3.0, 0.3, 0.1, 1e-3, -0.1], [3.0, 3.0, 0.9, 2.0, 1.3]

    def sampler(n, seed):
        return sample_theta(n, lo, hi, seed=seed)

    theta = sampler(64, 1)
    filt = {"mean_rt": 1.5, "choice_cnt": 10}
    lab = LanKDELabeller("angle", theta, n_samples=2000, n_kde=80, n_unif=10, n_neg=10, seed=3, launch=None,
                         filters=filt, theta_sampler=sampler)
    rows = lab.run()
    st = lab.filter_stats()
    assert lab.n_resampled > 0 and lab.kept.all()
    assert np.all(st[:, :, 1] <= 1.5) and np.all(st[:, :, 4] >= 10)
    out = lab.output_dict(rows)
    assert out["lan_data"].shape == (64 * 100, 7) and out["theta"].shape == (64, 5)
    assert np.array_equal(out["theta"], lab.theta)
    never = LanKDELabeller("angle", theta[:4], n_samples=500, n_kde=8, n_unif=1, n_neg=1, seed=4, launch=None,
                           filters={"mean_rt": -1.0}, theta_sampler=sampler, max_attempts=3)
    rows = never.run()
    assert not never.kept.any() and rows.shape[0] == 0
