"""GPU parity of the KDE fit / sample / eval kernels and the LAN row pipeline.

The oracle is the float64 LogKDE restatement (oracle/kde_ref.py, pinned to
scikit-learn). Tolerance: log-likelihoods within 2e-3 nats where the oracle
is above the floor + 1 (float32 accumulation over up to 1e5 samples), the
floor itself exactly; bandwidths within 1e-4 relative; counts exact.
"""

import math

import numpy as np
import pytest
from scipy import stats

from oracle import kde_ref

pytestmark = pytest.mark.gpu
F32 = np.float32
LB = -66.774


def _synthetic_sets(n_sets, n, seed=0, omit_frac=0.01):
    rng = np.random.default_rng(seed)
    rts, chs = [], []
    for k in range(n_sets):
        p = rng.uniform(0.05, 0.95)
        ch = np.where(rng.uniform(size=n) < p, 1, -1).astype(np.int32)
        rt = (rng.uniform(0.1, 1.5) + rng.lognormal(-0.5 + 0.2 * k % 3, 0.6, size=n)).astype(F32)
        rt[rng.uniform(size=n) < omit_frac] = -999.0
        rts.append(rt)
        chs.append(ch)
    return rts, chs


def _bank(rts, chs, choice_values=(-1, 1)):
    from ssmhip import _device as D
    from ssmhip.kde import KDEBank

    offs = np.concatenate([[0], np.cumsum([r.size for r in rts])]).astype(np.int64)
    bank = KDEBank(D.to_dev(np.concatenate(rts)), D.to_dev(np.concatenate(chs).astype(np.int8)), D.to_dev(offs),
                   list(choice_values))
    return bank, offs


def test_fit_matches_oracle(torch_gpu):
    rts, chs = _synthetic_sets(6, 20000, seed=1)
    rts[2][:] = -999.0  # a set with only omissions
    chs[3][:] = 1  # a set with one choice only
    rts[4] = rts[4][:1]  # a set with a single draw
    chs[4] = chs[4][:1]
    bank, _ = _bank(rts, chs)
    cnt, tot, h = bank.host_state()
    for k in range(6):
        fit = kde_ref.KDEFit(rts[k], chs[k], [-1, 1])
        for c in range(2):
            assert cnt[k, c] == fit.logs[c].size
            assert tot[k, c] == int(round(fit.p[c] * fit.N))
            if fit.h[c] is None:
                assert h[k, c] == 0.0
            else:
                assert abs(h[k, c] - fit.h[c]) <= 1e-4 * fit.h[c]


@pytest.mark.parametrize("vals", [(-1, 1), (0, 1, 2), (0, 1, 2, 3, 4, 5)])
def test_fit_compacts_each_choice_in_draw_order(torch_gpu, vals):
    """Each (set, choice) group holds log(rt) of that choice's kept draws in draw order, the
    data LogKDE keeps (kde_class.py:438-504), with the oracle's counts and bandwidths. Sets
    are empty, single, across the kernel's 4096-draw tiles and ragged; one set lacks a choice.
    2, 3 and 6 possible choices take the kernel's three per-choice register layouts."""
    print(f"[case] kde fit order, choices {vals}", flush=True)
    rng = np.random.default_rng(31 + len(vals))
    sizes = [0, 1, 4095, 4097, 8192, 12289, 30001]
    rts, chs = [], []
    for n in sizes:
        ch = rng.choice(np.array(vals), n).astype(np.int8)
        rt = (0.2 + rng.lognormal(-0.4, 0.7, n)).astype(F32)
        rt[rng.uniform(size=n) < 0.02] = -999.0
        rts.append(rt)
        chs.append(ch)
    chs[4][chs[4] == vals[-1]] = vals[0]
    bank, _ = _bank(rts, chs, vals)
    samples = bank.samples.cpu().numpy()
    start = bank.kde_start.cpu().numpy().reshape(len(sizes), len(vals))
    cnt, tot, h = bank.host_state()
    assert np.all(cnt[0] == 0) and np.all(tot[0] == 0) and np.all(h[0] == 0.0)
    for k in range(1, len(sizes)):
        fit = kde_ref.KDEFit(rts[k], chs[k], list(vals))
        for c in range(len(vals)):
            got = samples[start[k, c]:start[k, c] + cnt[k, c]]
            ref = fit.logs[c].astype(F32)
            assert got.shape == ref.shape
            # device logf against NumPy's float32 log: a few ulp; any reordering is far larger
            np.testing.assert_allclose(got, ref, rtol=5e-7, atol=5e-7)
            assert tot[k, c] == int(round(fit.p[c] * fit.N))
            if fit.h[c] is None:
                assert h[k, c] == 0.0
            else:
                assert abs(h[k, c] - fit.h[c]) <= 1e-4 * fit.h[c]


def test_eval_matches_oracle(torch_gpu):
    from ssmhip.kde import PointBatch

    rts, chs = _synthetic_sets(1, 100000, seed=2)
    bank, _ = _bank(rts, chs)
    fit = kde_ref.KDEFit(rts[0], chs[0], [-1, 1])
    rng = np.random.default_rng(3)
    q = np.concatenate([rng.uniform(0.05, 6, 600), [-0.5, 0.0, 5e-5, 19.9, 1e-4, 0.3]]).astype(F32)
    qc = rng.integers(0, 2, q.size)
    pts, order = PointBatch.from_host(q, qc, 1, 2)
    ll = bank.evaluate(pts).cpu().numpy()
    got = np.empty_like(ll)
    got[order] = ll
    ref = fit.eval(q.astype(np.float64), np.array([-1, 1])[qc])
    above = ref > LB + 1
    assert np.max(np.abs(got[above] - ref[above])) < 2e-3
    assert np.all(got[~above] <= LB + 1 + 2e-3)
    assert np.all(got[ref == LB] == np.float32(LB))


def test_eval_no_base_data_and_single_choice(torch_gpu):
    from ssmhip.kde import PointBatch

    rts = [np.full(500, 0.8, F32) + np.linspace(0, 0.5, 500, dtype=F32)]
    chs = [np.ones(500, np.int32)]
    bank, _ = _bank(rts, chs)
    pts, order = PointBatch.from_host(np.array([0.9, 1.0], F32), np.array([0, 1]), 1, 2)
    ll = bank.evaluate(pts, max_t=20.0).cpu().numpy()[np.argsort(order)]
    assert abs(ll[0] - (-math.log(500 * 20.0))) < 1e-5  # choice -1 has no base data
    assert ll[1] > LB


def test_sample_counts_and_distribution(torch_gpu):
    rts, chs = _synthetic_sets(3, 50000, seed=4)
    bank, _ = _bank(rts, chs)
    pts = bank.sample(800, 200, max_t=20.0, seed=9)
    rt = pts.rt.cpu().numpy().reshape(3, 1000)
    ch = pts.choice.cpu().numpy().reshape(3, 1000)
    gb = pts.grp_begin.cpu().numpy().reshape(3, 2)
    gc = pts.grp_count.cpu().numpy().reshape(3, 2)
    cnt, tot, h = bank.host_state()
    for k in range(3):
        assert gc[k].sum() == 1000 and gb[k, 0] == k * 1000 and gb[k, 1] == k * 1000 + gc[k, 0]
        assert np.all(ch[k][: gc[k, 0]] == 0) and np.all(ch[k][gc[k, 0]:] == 1)
        assert gc[k].sum() - 200 == 800
        assert np.all(rt[k] > 0)
    # KDE resamples follow the KDE: compare log rt of a large resample with the oracle's KDE draws
    big = bank.sample(1000, 0, seed=11)
    r = big.rt.cpu().numpy().reshape(3, 1000)
    c = big.choice.cpu().numpy().reshape(3, 1000)
    rng = np.random.default_rng(5)
    fit = kde_ref.KDEFit(rts[0], chs[0], [-1, 1])
    for ci in range(2):
        mine = np.log(r[0][c[0] == ci])
        if mine.size < 100:
            continue
        ref = rng.choice(fit.logs[ci], 20000) + fit.h[ci] * rng.standard_normal(20000)
        assert stats.ks_2samp(mine, ref).pvalue > 0.001


def test_sample_counts_follow_the_reference_rule(torch_gpu):
    # kde_class.py:387-399 (oracle kde_ref.n_by_choice): n_c = round(800 p_c) on the float64
    # product, +-1 fix-up. Sets of 320 draws with 23, 49, 51 or 87 draws of choice -1 are ties in
    # rationals where the float64 product rounds the other way; a set of 1600 draws with one -1.
    rng = np.random.default_rng(21)
    rts, chs = _synthetic_sets(4, 30000, seed=22)
    for n_minus, N in ((23, 320), (49, 320), (51, 320), (87, 320), (1, 1600), (7, 7), (0, 500)):
        ch = np.ones(N, np.int32)
        ch[rng.permutation(N)[:n_minus]] = -1
        rts.append(rng.uniform(0.3, 2.0, N).astype(F32))
        chs.append(ch)
    bank, _ = _bank(rts, chs)
    K = len(rts)
    pts = bank.sample(800, 0, seed=3)
    gc = pts.grp_count.cpu().numpy().reshape(K, 2)
    for k in range(K):
        fit = kde_ref.KDEFit(rts[k], chs[k], [-1, 1])
        assert list(gc[k]) == fit.n_by_choice(800), (k, list(gc[k]))
    assert list(gc[4]) == [57, 743]  # the rational rule would give 58 + 742


def test_logkde_dropin_contract(torch_gpu):
    from ssmhip import simulator
    from ssmhip.kde import LogKDE

    data = simulator.Simulator("ddm").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3}, n_samples=1000,
                                                 random_state=3)
    kde = LogKDE(simulator_data=data)
    assert kde.simulator_info is data["metadata"] or kde.simulator_info == data["metadata"]
    assert kde.auto_bandwidth is True and kde.bandwidth_type == "silverman"
    assert len(kde.compute_bandwidths()) == 2
    res = kde.kde_eval({"rts": np.array([0.6, 0.8]), "choices": np.array([1, -1])}, lb=-66.774)
    assert isinstance(res, np.ndarray) and len(res) == 2 and np.all(res >= -66.774)
    fit = kde_ref.KDEFit(np.asarray(data["rts"]).reshape(-1), np.asarray(data["choices"]).reshape(-1), [-1, 1])
    ref = fit.eval(np.array([0.6, 0.8]), np.array([1, -1]))
    assert np.max(np.abs(res - ref)) < 2e-3
    s = kde.kde_sample(n_samples=100)
    assert len(s["rts"]) == 100 and len(s["choices"]) == 100
    assert set(np.unique(s["choices"])) <= {-1.0, 1.0}
    s2 = kde.kde_sample(n_samples=3000, use_empirical_choice_p=False, alternate_choice_p=np.array([0.25, 0.75]))
    assert abs(np.mean(s2["choices"] == 1) - 0.75) < 1e-3


def test_lan_rows_small_batch_against_oracle(torch_gpu):
    from ssmhip.lan import LanKDELabeller, sample_angle_uniform

    theta = sample_angle_uniform(6, seed=7)
    lab = LanKDELabeller("angle", theta, n_samples=20000, launch=None, seed=5)
    rows = lab.run().cpu().numpy()
    assert rows.shape == (6 * 1000, 8)
    sim_rts = lab.sim.rts.cpu().numpy()
    sim_ch = lab.sim.choices.cpu().numpy()
    offs = lab.sim.offsets
    neg_rt, neg_ch = [], []
    for k in range(6):
        blk = rows[k * 1000:(k + 1) * 1000]
        assert np.allclose(blk[:, :5], theta[k][None, :])
        # the reference mixture: 100 negative-RT rows per set with the fixed label
        neg = blk[:, 5] < 1e-4
        assert neg.sum() == 100
        assert np.all(blk[neg, 7] == np.float32(-66.77497)) and np.all(blk[neg, 5] >= -1.0)
        assert set(np.unique(blk[neg, 6])) <= {-1.0, 1.0}
        # shuffled together with the other rows: their positions spread over the block
        assert 350 < np.flatnonzero(neg).mean() < 650
        neg_rt.append(blk[neg, 5])
        neg_ch.append(blk[neg, 6])
        pos = blk[~neg]
        fit = kde_ref.KDEFit(sim_rts[offs[k]:offs[k + 1]], sim_ch[offs[k]:offs[k + 1]], [-1, 1])
        ref = fit.eval(pos[:, 5].astype(np.float64), pos[:, 6])
        above = ref > LB + 1
        assert np.max(np.abs(pos[above, 7] - ref[above])) < 2e-3
        assert np.all(pos[:, 7] >= LB)  # the KDE floor; the negative-RT label lies just below it
        assert set(np.unique(blk[:, 6])) <= {-1.0, 1.0}
        # 100 uniform rows in [1e-4, 20] plus 800 KDE rows
        assert np.sum(~neg) == 900 and np.all(pos[:, 5] >= 1e-4) and np.all(pos[:, 5] <= 20.0)
    # negative rows: rt ~ U(-1, 1e-4), both choices
    assert stats.kstest(np.concatenate(neg_rt), stats.uniform(loc=-1.0, scale=1.0001).cdf).pvalue > 1e-3
    assert 0.4 < np.mean(np.concatenate(neg_ch) == 1.0) < 0.6


def test_lan_rows_shuffle_is_a_permutation(torch_gpu):
    from ssmhip.lan import LanKDELabeller, sample_angle_uniform

    theta = sample_angle_uniform(3, seed=8)
    a = LanKDELabeller("angle", theta, n_samples=5000, launch=None, seed=2, shuffle=False)
    ra = a.run().cpu().numpy()
    b = LanKDELabeller("angle", theta, n_samples=5000, launch=None, seed=2, shuffle=True)
    rb = b.run().cpu().numpy()
    for k in range(3):
        x = ra[k * 1000:(k + 1) * 1000]
        y = rb[k * 1000:(k + 1) * 1000]
        assert not np.array_equal(x, y)
        assert np.array_equal(x[np.lexsort(x.T)], y[np.lexsort(y.T)])
        assert np.all(x[900:, 5] < 1e-4) and np.all(x[:900, 5] >= 1e-4)  # unshuffled: negative rows last
