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
