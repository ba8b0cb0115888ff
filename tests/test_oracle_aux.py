"""Pins of the auxiliary-label oracle on a hand-built case (counts written out by hand)."""

import numpy as np

from oracle import aux_ref


def test_hand_built_set():
    # max_t = 20: 128 bins of 0.15625 s, 256 bins of 0.078125 s
    rts = np.array([0.0, 0.15625, 0.156, 19.99, 20.0, 20.5, -999.0, 1.0, 0.078125, 3.0], np.float32)
    ch = np.array([1, 1, -1, -1, 1, 1, 1, -1, 1, 0], np.int32)
    d = aux_ref.labels_one_set(rts, ch, [-1, 1], 20.0)
    # choice +1: rts 0.0, 0.15625, 20.0, 20.5 (above max_t: dropped), -999 (omitted), 0.078125
    h = d["binned_128"][0, :, 1]
    assert h[0] == 2 and h[1] == 1 and h[127] == 1 and h.sum() == 4  # 0.0, 0.078125 | 0.15625 | 20.0
    h = d["binned_256"][0, :, 1]
    assert h[0] == 1 and h[1] == 1 and h[2] == 1 and h[255] == 1 and h.sum() == 4
    # choice -1: rts 0.156 (bin 0 of 128: below the 0.15625 edge in float32 -> float64), 19.99, 1.0
    h = d["binned_128"][0, :, 0]
    assert h[0] == 1 and h[127] == 1 and h[6] == 1 and h.sum() == 3  # 1.0 / 0.15625 = 6.4
    assert np.allclose(d["choice_p"], [[3 / 10, 6 / 10]])
    assert np.allclose(d["choice_p_no_omission"], d["choice_p"])  # choices never carry -999
    assert d["omission_p"][0, 0] == 0.0
    # no-go: choice != 1 (3 of -1, one 0) or rt == -999 (one, with choice 1) -> 5 of 10
    assert d["nogo_p"][0, 0] == 0.5 and d["go_p"][0, 0] == 0.5


def test_concatenation_shapes_and_float32():
    rng = np.random.default_rng(0)
    rts = rng.uniform(0, 25, 300).astype(np.float32)
    ch = rng.choice([0, 1, 2], 300).astype(np.int32)
    offs = np.array([0, 100, 180, 300])
    d = aux_ref.labels_all_sets(rts, ch, offs, [0, 1, 2])
    assert d["cpn_labels"].shape == (3, 3) and d["cpn_labels"].dtype == np.float32
    assert d["binned_256"].shape == (3, 256, 3) and d["gonogo_labels"].shape == (3, 1)
    two = aux_ref.labels_all_sets(rts, np.where(ch > 0, 1, -1), offs, [-1, 1])
    assert two["cpn_labels"].shape == (3, 1)
    assert np.isclose(two["cpn_labels"][0, 0], np.mean(np.where(ch[:100] > 0, 1, -1) == 1))


def test_filter_stats_hand_case():
    """Mode ties resolve to the smaller value, omissions and foreign choices are excluded, a
    choice with under 5 values has mode share 0, an empty choice gets (-1, -1, 1, 0, 0)."""
    rts = np.array([0.5, 0.7, 0.7, 0.5, 0.9, -999.0, 0.3, 0.3, 0.3, 0.8, 0.8, 1.0], np.float32)
    ch = np.array([1, 1, 1, 1, 1, 1, -1, -1, 5, 1, 1, 1])
    st = aux_ref.filter_stats_one_set(rts, ch, [-1, 1, 2])
    # choice -1: [0.3, 0.3] -> mode 0.3 (count 2), n = 2 < 5 -> share 0
    assert st[0, 0] == np.float32(0.3) and st[0, 3] == 0 and st[0, 4] == 2
    # choice 1: [0.5, 0.7, 0.7, 0.5, 0.9, 0.8, 0.8, 1.0]: 0.5, 0.7, 0.8 twice each -> 0.5, share 2/8
    x = np.array([0.5, 0.7, 0.7, 0.5, 0.9, 0.8, 0.8, 1.0], np.float32)
    assert st[1, 0] == np.float32(0.5) and st[1, 3] == 0.25 and st[1, 4] == 8
    assert abs(st[1, 1] - x.mean()) < 1e-7 and abs(st[1, 2] - x.std()) < 1e-7
    assert tuple(st[2]) == (-1, -1, 1, 0, 0)
    assert aux_ref.filter_keep(st, {})
    assert not aux_ref.filter_keep(st, {"choice_cnt": 1})  # choice 2 has no draws
    assert not aux_ref.filter_keep(st, {"mean_rt": 0.5}) and aux_ref.filter_keep(st, {"mean_rt": 0.8})
    assert not aux_ref.filter_keep(st, {"mode_cnt_rel": 0.2}) and aux_ref.filter_keep(st, {"mode_cnt_rel": 0.25})
