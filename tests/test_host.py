"""Host-side logic that needs no GPU: preprocessing, records, configs, strategies."""

import numpy as np
import pytest

F32 = np.float32


def test_records_carry_the_reference_per_trial_float32_quantities():
    from oracle import em_numpy as E
    from ssmhip import _lib as L
    from ssmhip import em

    v, a, z, t = (np.array(x, F32) for x in ([0.5, -1.2], [1.5, 0.7], [0.4, 0.6], [0.3, 1.1]))
    th = np.array([0.2, -0.05], F32)
    rec = em.ddm_records(v, a, z, t, np.array([999.0, 1.0], F32), np.ones(2, F32), 20.0, 0.001, L.BOUNDARY_ANGLE,
                         {"a": a, "theta": th})
    assert np.array_equal(rec[:, L.P_DEADLINE_TMP], np.minimum(F32(20.0), (np.array([999.0, 1.0], F32) - t)))
    assert rec[0, L.P_SQRT_ST] == E.sqrt_dt32(0.001)
    # the slope the kernel uses reproduces NumPy's float32 boundary exactly
    table = E.boundary_table("angle", {"a": a, "theta": th}, 2, 20.0, 0.001)
    ix = np.arange(table.shape[1], dtype=F32)
    for k in range(2):
        b = (a[k] + (ix * F32(0.001)).astype(F32) * rec[k, L.P_B1]).astype(F32)
        assert np.array_equal(b, table[k])


def test_kernel_boundary_codes():
    from ssmhip import _lib as L
    from ssmhip import boundary_functions as BF

    assert BF.kernel_boundary_code(BF.constant) == L.BOUNDARY_CONSTANT
    assert BF.kernel_boundary_code(BF.angle) == L.BOUNDARY_ANGLE
    assert BF.kernel_boundary_code(BF.weibull_cdf) == L.BOUNDARY_WEIBULL

    def angle(t, a, theta):  # a user function that merely shares the name
        return a - t * theta

    assert BF.kernel_boundary_code(angle) == L.BOUNDARY_TABLE
    assert BF.n_steps(20.0, 0.001) == 20001


def test_theta_preprocessing_matches_reference_rules():
    from ssmhip.simulator import _make_valid_dict, _preprocess_theta, _theta_array_to_dict, make_noise_vec

    d = _make_valid_dict({"v": 1.0, "a": [1.0, 2.0], "z": np.array([0.5, 0.4])})
    assert all(x.dtype == F32 and x.shape == (2,) for x in d.values())
    with pytest.raises(ValueError):
        _make_valid_dict({"v": [1.0, 2.0], "a": [1.0, 2.0, 3.0]})
    arr = _preprocess_theta([[0.5, 1.0, 0.5, 0.3]])
    assert arr.dtype == F32
    with pytest.raises(ValueError):
        _theta_array_to_dict(arr, ["v", "a", "z"])
    assert make_noise_vec(1.0, 3, 1).shape == (3,)
    nv = make_noise_vec(0.1, 3, 4)
    assert nv.shape == (3, 4) and np.all(nv == F32(0.1))


def test_model_configs_and_deadline_variant():
    from ssmhip import config

    for name in ("ddm", "angle", "weibull", "race_4", "lca_4", "lba2", "lba_angle_3"):
        c = config.get(name)
        assert c["n_params"] == len(c["params"]) == len(c["param_bounds"][0]) == len(c["default_params"])
    c = config.get("angle_deadline")
    assert c["params"][-1] == "deadline" and c["name"] == "angle_deadline"
    with pytest.raises(ValueError):
        config.get("nonexistent_model")


def test_race_and_lca_variant_configs():
    """race.py / lca.py variants: z per accumulator, one shared z (no_bias) or zeros (no_z),
    constant or angle boundary; records carry the trial's own a and the angle slope."""
    from ssmhip import config, em
    from ssmhip import _lib as L
    from ssmhip.boundary_functions import kernel_boundary_code
    from ssmhip.simulator import make_boundary_dict

    names = [m for m in config.list_models() if m.startswith(("race_", "lca_"))]
    assert len(names) == 25
    rng = np.random.default_rng(3)
    for name in names:
        c = config.get(name)
        n = c["n_particles"]
        assert c["n_params"] == len(c["params"]) == len(c["param_bounds"][0]) == len(c["default_params"])
        assert c["choices"] == list(range(n)) and c["boundary_name"] == ("angle" if "angle" in name else "constant")
        lo, hi = (np.asarray(b, F32) for b in c["param_bounds"])
        rows = lo + (hi - lo) * rng.uniform(size=(3, lo.size)).astype(F32)
        theta = {p: rows[:, j].copy() for j, p in enumerate(c["params"])}
        theta["s"] = np.ones((3, n), F32)
        theta["deadline"] = np.full(3, 999, F32)
        for tr in c["simulation_transforms"]:
            theta = tr(theta, c, 3)
        assert theta["z"].shape == (3, n) and theta["v"].shape == (3, n)
        if "no_z" in name:
            assert np.all(theta["z"] == 0)
        elif "no_bias" in name:
            assert np.all(theta["z"] == rows[:, c["params"].index("z")][:, None])
        else:
            assert np.array_equal(theta["z"], rows[:, [c["params"].index(f"z{i}") for i in range(n)]])
        bd = make_boundary_dict(c, theta)
        code = kernel_boundary_code(bd["boundary_fun"])
        kw = {"g": theta["g"], "b": theta["b"]} if name.startswith("lca") else {}
        rec = em.accumulator_records(theta["v"], theta["z"], theta["t"], theta["s"], theta["deadline"], 20.0, 1e-3,
                                     code, bd["boundary_params"], **kw)
        assert np.array_equal(rec[:, L.P_A], rows[:, c["params"].index("a")])
        if "angle" in name:
            assert code == L.BOUNDARY_ANGLE
            th = rows[:, c["params"].index("theta")]
            assert np.allclose(rec[:, L.P_B1], -np.tan(th.astype(np.float64)), rtol=1e-5, atol=1e-6)
    assert config.get("weibull_cdf")["params"] == config.get("weibull")["params"]
    assert config.get("full_ddm2")["simulator"] is config.get("full_ddm")["simulator"]


def test_compact_single_particle_records():
    """ddm_records(stride=16) holds exactly the first 16 floats of the full records."""
    from ssmhip import _lib as L
    from ssmhip import em

    rng = np.random.default_rng(12)
    n = 9
    args = [rng.uniform(0.1, 0.9, n).astype(F32) for _ in range(4)] + [np.full(n, 999, F32), np.ones(n, F32)]
    bp = {"a": rng.uniform(0.5, 2, n).astype(F32), "theta": rng.uniform(-0.1, 1.3, n).astype(F32)}
    full = em.ddm_records(*args, 20.0, 1e-3, L.BOUNDARY_ANGLE, bp, alpha=np.full(n, 1.5, F32))
    compact = em.ddm_records(*args, 20.0, 1e-3, L.BOUNDARY_ANGLE, bp, alpha=np.full(n, 1.5, F32),
                             stride=L.SET_STRIDE_SINGLE)
    assert full.shape == (n, L.SET_STRIDE) and compact.shape == (n, L.SET_STRIDE_SINGLE)
    assert np.array_equal(full[:, :16], compact) and not np.any(full[:, 16:])
    with pytest.raises(ValueError):
        em.ddm_records(*args, 20.0, 1e-3, L.BOUNDARY_ANGLE, bp, stride=32)


def test_batched_records_shape_checks():
    from ssmhip import batched

    theta, ns = batched.sample_angle_ragged(8, seed=0)
    assert theta.shape == (8, 5) and np.all((ns >= 1000) & (ns <= 100000))
    with pytest.raises(ValueError):
        batched.records_for("angle", theta[:, :4], 0.001, 20.0)


def test_resample_mixture_strategy_with_a_stub_estimator():
    from ssmhip.estimators import ResampleMixtureStrategy

    class Stub:
        def get_metadata(self):
            return {"possible_choices": [-1, 1], "max_t": 20.0}

        def sample(self, n_samples, random_state=None):
            return {"rts": np.linspace(0.3, 2, n_samples), "choices": np.ones(n_samples)}

        def evaluate(self, rts, choices):
            return -np.asarray(rts)

    gen = {"training": {"n_samples_per_param": 1000, "mixture_probabilities": [0.8, 0.1, 0.1],
                        "separate_response_channels": False, "negative_rt_log_likelihood": -66.77497}}
    mc = {"params": ["v", "a", "z", "t"], "nchoices": 2}
    out = ResampleMixtureStrategy(gen, mc).generate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3}, Stub(), random_state=0)
    assert out.shape == (1000, 7) and out.dtype == F32
    assert np.sum(out[:, -1] == F32(-66.77497)) == 100
    assert np.all(out[:, 0] == 1.0)


def test_training_data_file_format(tmp_path):
    import pickle
    import re

    from ssmhip.lan import save_training_data

    data = {"lan_data": np.zeros((4, 7), np.float32), "lan_labels": np.ones(4, np.float32),
            "binned_128": np.zeros((1, 128, 2), np.float32)}
    path = save_training_data(data, tmp_path / "out", generator_config={"output": {"pickle_protocol": 4}},
                              model_config={"name": "angle"})
    assert re.fullmatch(r"training_data_[0-9a-f]{32}\.pickle", path.name)
    with path.open("rb") as f:
        assert f.read(2) == b"\x80\x04"  # protocol 4
    with path.open("rb") as f:
        back = pickle.load(f)  # a file this test wrote
    assert set(back) == {"lan_data", "lan_labels", "binned_128", "generator_config", "model_config"}
    assert np.array_equal(back["lan_data"], data["lan_data"]) and back["model_config"]["name"] == "angle"


def test_angle_slope_array_call_equals_the_per_trial_scalar_calls():
    """The reference evaluates the angle boundary with one scalar theta per trial; the records
    take the slope from one array call, which must give the same float32 values."""
    from ssmhip import em

    th = np.random.default_rng(0).uniform(-0.1, 1.3, 50_000).astype(np.float32)
    th[:3] = np.float32([0.0, 1.3, -0.1])
    scalar = np.array([(-np.sin(x) / np.cos(x)) for x in th], dtype=np.float32)
    assert np.array_equal(em.angle_slope32(th), scalar)
    # long inputs made of runs (a parameter broadcast over trials) are evaluated once per run:
    # runs of unequal lengths, a run of one, and the same value in two separate runs
    lens = np.random.default_rng(1).integers(1, 3000, 200)
    lens[7] = 1
    vals = th[:200].copy()
    vals[150] = vals[10]
    runs = np.repeat(vals, lens)
    assert runs.size >= 65536
    want = scalar[:200].copy()
    want[150] = scalar[10]
    assert np.array_equal(em.angle_slope32(runs), np.repeat(want, lens))


def test_deadline_tmp_is_the_float32_minimum_of_max_t_and_deadline_minus_t():
    """compute_deadline_tmp (_utils.pyx:712-718) per trial, in float32."""
    from ssmhip import em

    rng = np.random.default_rng(3)
    t = rng.uniform(0, 2, 1000).astype(np.float32)
    dl = np.where(rng.uniform(size=1000) < 0.5, 999.0, rng.uniform(0, 30, 1000)).astype(np.float32)
    want = np.array([min(np.float32(20.0), np.float32(d - x)) for d, x in zip(dl, t)], dtype=np.float32)
    got = em.deadline_tmp32(20.0, dl, t)
    assert got.dtype == np.float32 and np.array_equal(got, want)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says why, before any
    device work."""
    import subprocess
    import sys
    from pathlib import Path

    bench = Path(__file__).resolve().parents[1] / "bench.py"
    r = subprocess.run([sys.executable, str(bench), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""


def test_autotune_cache_key_separates_batch_sizes():
    from types import SimpleNamespace

    from ssmhip import autotune

    def job(n_sets, n_total, n_steps=20001):
        return SimpleNamespace(prob=SimpleNamespace(model=0, boundary=1, n_particles=1, replay_stride=0,
                                                    n_sets=n_sets, n_total=n_total, n_steps=n_steps))

    small, big = autotune._key(job(64, 1_300_000)), autotune._key(job(4096, 86_000_000))
    assert small != big
    assert autotune._key(job(64, 1_300_000)) == autotune._key(job(100, 1_500_000))  # same size classes
    assert autotune._key(job(64, 1_300_000, n_steps=5001)) != small
