"""HSSM support: shape, seed and validation logic on the CPU (ssms/hssm_support.py), and the
posterior-predictive draw through the GPU simulator (marked gpu)."""

import numpy as np
import pytest


def _hs():
    from ssmhip import hssm_support as H

    return H


def test_prepare_theta_scalar_parameters_tile_one_row():
    H = _hs()
    args = [np.array([0.5]), np.array([1.0]), np.array([0.4]), np.array([0.2])]
    scalar, theta, max_shape, n = H._prepare_theta_and_shape(args, 7)
    assert scalar and theta.shape == (7, 4) and max_shape is None and n is None
    assert np.all(theta[3] == [0.5, 1.0, 0.4, 0.2])
    assert H._prepare_theta_and_shape(args, (3,))[1].shape == (3, 4)
    with pytest.raises(ValueError, match="Size must be"):
        H._prepare_theta_and_shape(args, None)


def test_prepare_theta_vector_parameters_broadcast_to_the_largest():
    H = _hs()
    v = np.arange(6.0).reshape(2, 3)  # 2 draws x 3 data trials
    a = np.array([1.0, 1.5, 2.0])     # per trial
    scalar, theta, max_shape, n = H._prepare_theta_and_shape([v, a, np.array(0.5)], 12)
    assert not scalar and max_shape == (2, 3) and n == 3
    assert theta.shape == (6, 3)
    assert np.array_equal(theta[:, 0], np.arange(6.0)) and np.array_equal(theta[:, 1], np.tile(a, 2))
    assert np.all(theta[:, 2] == 0.5)


def test_only_an_owned_float32_theta_is_used_without_a_copy():
    """The reference copies theta to float32 before simulating (simulator.py:221-254); rng_fn's
    own per-call matrix, marked OwnedTheta, is used as it is, and every other array is still
    copied, so a caller's array never aliases the simulator's input."""
    from ssmhip.simulator import OwnedTheta, _preprocess_theta

    theta = np.asfortranarray(np.arange(12, dtype=np.float32).reshape(4, 3))
    copied = _preprocess_theta(theta)
    assert not np.shares_memory(copied, theta) and np.array_equal(copied, theta)
    owned = _preprocess_theta(theta.view(OwnedTheta))
    assert np.shares_memory(owned, theta) and type(owned) is np.ndarray and owned.dtype == np.float32
    theta64 = theta.astype(np.float64).view(OwnedTheta)
    out = _preprocess_theta(theta64)
    assert out.dtype == np.float32 and not np.shares_memory(out, theta64)
    H = _hs()
    assert getattr(H.get_simulator_fun_internal("angle"), "takes_owned_theta", False)
    assert not getattr(H.get_simulator_fun_internal(lambda **kw: kw), "takes_owned_theta", False)


def test_replicas_sizes_and_output_shape():
    H = _hs()
    assert H._calculate_n_replicas(True, 10, None) == 1
    assert H._calculate_n_replicas(False, None, 3) == 1 and H._calculate_n_replicas(False, 1, 3) == 1
    assert H._calculate_n_replicas(False, 12, 3) == 4 and H._calculate_n_replicas(False, (12,), 3) == 4
    with pytest.raises(ValueError, match="multiple"):
        H._calculate_n_replicas(False, 10, 3)
    assert H._reshape_sims_out((2, 3), 1, 2) == (2, 3, 2)
    assert H._reshape_sims_out((2, 3), 4, 2) == (2, 3, 4, 2)


def test_seed_range_and_extra_fields_alignment():
    H = _hs()
    rng = np.random.default_rng(0)
    seeds = [H._get_seed(rng) for _ in range(2000)]
    assert min(seeds) >= 0 and max(seeds) < 2**31
    ef = H._broadcast_extra_fields({"cov": np.array([1.0, 2.0, 3.0]), "k": np.array(5.0),
                                    "other": np.arange(4.0)}, (2, 3), 3)
    assert np.array_equal(ef["cov"], [1, 2, 3, 1, 2, 3]) and ef["k"] == 5.0 and ef["other"].shape == (4,)


def test_simulator_fun_resolution_and_validation(caplog):
    H = _hs()
    f = H.get_simulator_fun_internal("angle")
    assert H.validate_simulator_fun(f) == ("angle", [-1, 1], 2)
    assert H.get_simulator_fun_internal("race_3").choices == [0, 1, 2]
    with caplog.at_level("WARNING"):
        g = H.get_simulator_fun_internal("no_such_model")
    assert "not supported" in caplog.text and g.choices == [0, 1, 2]

    def custom(**kw):
        return None

    assert H.get_simulator_fun_internal(custom) is custom
    with pytest.raises(ValueError, match="model_name"):
        H.validate_simulator_fun(custom)
    with pytest.raises(ValueError, match="string or a callable"):
        H.get_simulator_fun_internal(3)
    deco = H.decorate_atomic_simulator("m", obs_dim=2)(lambda: None)
    assert (deco.model_name, deco.choices, deco.obs_dim) == ("m", [-1, 1], 2)
    deco.obs_dim = 2.0
    with pytest.raises(ValueError, match="integer"):
        H.validate_simulator_fun(deco)


@pytest.mark.gpu
def test_rng_fn_posterior_predictive_shapes_and_per_trial_parameters(torch_gpu):
    """Vector parameters (2 posterior draws x 4000 data trials) and 3 replicas: one simulator
    call over 8000 parameter rows; each trial keeps its own parameters (drift sign -> choice)."""
    H = _hs()
    fn = H.get_simulator_fun_internal("ddm")
    n_trials = 4000
    v = np.where(np.arange(n_trials) % 2 == 0, 3.0, -3.0)[None, :].repeat(2, axis=0)
    args = [v, np.array(1.0), np.array(0.5), np.array(0.3)]
    out = H.rng_fn(args, 3 * n_trials, np.random.default_rng(1), fn, 2)
    assert out.shape == (2, n_trials, 3, 2)
    single = H.rng_fn(args, None, np.random.default_rng(2), fn, 2)
    assert single.shape == (2, n_trials, 2)
    rt, ch = single[..., 0], single[..., 1]
    assert np.all(rt > 0.3)
    assert np.mean(ch[:, 0::2] == 1) > 0.97 and np.mean(ch[:, 1::2] == -1) > 0.97
    scalar = H.rng_fn([np.array(1.0), np.array(1.2), np.array(0.5), np.array(0.2)], 500, np.random.default_rng(3),
                      fn, 2)
    assert scalar.shape == (500, 2)


@pytest.mark.gpu
def test_rng_fn_per_trial_draws_match_cpu_restatement_statistically(torch_gpu):
    """Angle-model posterior-predictive draws with per-trial parameters: 2 posterior draws x
    40,000 trials, trial i using the i mod 4-th of four (v, theta) pairs. The draws pooled per
    pair (20,000 each) against the C++/OpenMP restatement at 1e5 samples of that pair: choice
    probabilities, KS p > 0.01 and the 0.1..0.9 rt quantiles within 3 MC SE."""
    from em_cases import compare_samples
    from oracle import cpu_baseline
    from oracle import em_numpy as E

    H = _hs()
    f32 = np.float32
    vs = np.array([1.2, -0.5, 0.2, 2.0], f32)
    ths = np.array([0.1, 0.6, -0.05, 1.0], f32)
    a, z, t = f32(1.3), f32(0.45), f32(0.25)
    n_trials = 40_000
    pair = np.arange(n_trials) % 4
    v = np.broadcast_to(vs[pair], (2, n_trials)).copy()
    th = np.broadcast_to(ths[pair], (2, n_trials)).copy()
    fn = H.get_simulator_fun_internal("angle")
    out = H.rng_fn([v, np.array(a), np.array(z), np.array(t), th], None, np.random.default_rng(8), fn, 2)
    assert out.shape == (2, n_trials, 2)
    N = 100_000
    bp = {"a": np.full(4, a, f32), "theta": ths}
    table = E.boundary_table("angle", bp, 4, 20.0, 0.001)
    rc, cc, _ = cpu_baseline.em_simulate(0, vs, np.full(4, z, f32), np.full(4, E.sqrt_dt32(0.001)), np.full(4, t, f32),
                                         np.full(4, 999.0, f32), np.full(4, 20.0, f32), table,
                                         np.arange(5, dtype=np.int64) * N, dt=0.001, seed=5, smooth_unif=True)
    for j in range(4):
        sel = out[:, pair == j, :].reshape(-1, 2)
        compare_samples(sel[:, 0].astype(f32), sel[:, 1].astype(np.int32), rc[j * N:(j + 1) * N],
                        cc[j * N:(j + 1) * N], (-1, 1))
