"""The TrainingDataGenerator drop-in: configuration handling on the CPU, the device batch and
its training_data_<uuid>.pickle on the GPU."""

import pickle

import numpy as np
import pytest


def test_mixture_counts_follow_the_reference_rounding():
    from ssmhip.training_data import mixture_counts

    assert mixture_counts(1000, [0.8, 0.1, 0.1]) == (800, 100, 100)
    assert mixture_counts(999, [0.8, 0.1, 0.1]) == (801, 99, 99)  # remainder to the KDE rows
    assert mixture_counts(10, [0.8, 0.2]) == (8, 2, 0)
    with pytest.raises(ValueError, match="Rounding"):
        mixture_counts(10, [0.5, 0.6, 0.5])


def test_generator_config_and_model_checks():
    from ssmhip.training_data import TrainingDataGenerator, get_lan_kde_config

    cfg = get_lan_kde_config("angle")
    assert set(cfg) == {"pipeline", "estimator", "training", "simulator", "output", "model"}
    gen = TrainingDataGenerator(cfg, seed=1)
    assert gen.model_config["name"] == "angle" and gen.generator_config["estimator"]["displace_t"] is False
    with pytest.raises(NotImplementedError, match="only 'kde'"):
        TrainingDataGenerator({**cfg, "estimator": {"type": "pyddm"}})
    with pytest.raises(NotImplementedError, match="no batched device path"):
        TrainingDataGenerator(get_lan_kde_config("lba3"))


@pytest.mark.gpu
def test_generate_data_training_writes_the_reference_file(torch_gpu, tmp_path):
    from ssmhip.training_data import TrainingDataGenerator, get_lan_kde_config

    cfg = get_lan_kde_config("ddm")
    cfg["pipeline"]["n_parameter_sets"] = 32
    cfg["simulator"]["n_samples"] = 5000
    cfg["training"]["n_samples_per_param"] = 200
    cfg["output"]["folder"] = str(tmp_path)
    cfg["simulator"]["filters"] = {"mean_rt": 10.0, "choice_cnt": 1}
    data = TrainingDataGenerator(cfg, seed=11).generate_data_training(save=True)
    files = list(tmp_path.glob("training_data_*.pickle"))
    assert len(files) == 1
    with open(files[0], "rb") as f:  # written by this test
        loaded = pickle.load(f)
    for key in ("lan_data", "lan_labels", "cpn_data", "cpn_labels", "cpn_no_omission_data", "cpn_no_omission_labels",
                "opn_data", "opn_labels", "gonogo_data", "gonogo_labels", "binned_128", "binned_256", "theta",
                "generator_config", "model_config"):
        assert key in loaded, key
    assert loaded["lan_data"].shape == (32 * 200, 6) and loaded["lan_labels"].shape == (32 * 200,)
    assert loaded["lan_data"].dtype == np.float32 and loaded["binned_128"].shape == (32, 128, 2)
    lo, hi = (np.array(b, np.float32) for b in loaded["model_config"]["param_bounds"])
    assert np.all((data["theta"] >= lo) & (data["theta"] <= hi))
    assert np.sum(loaded["lan_labels"] == np.float32(-66.77497)) == 32 * 20  # the negative-rt rows


def test_yaml_configuration_maps_like_the_reference_cli(tmp_path):
    from ssmhip.generate import collect_data_generator_config

    yml = {"MODEL": "angle", "GENERATOR_APPROACH": "lan", "PIPELINE": {"N_PARAMETER_SETS": 100, "N_SUBRUNS": 20},
           "SIMULATOR": {"N_SAMPLES": 2000, "DELTA_T": 0.001}, "TRAINING": {"N_SAMPLES_PER_PARAM": 200},
           "ESTIMATOR": {"TYPE": "KDE"}}
    cfg = collect_data_generator_config(yml, tmp_path)
    d = cfg["data_config"]
    assert cfg["model_config"]["name"] == "angle" and d["model"] == "angle"
    assert d["pipeline"]["n_parameter_sets"] == 100 and d["pipeline"]["n_subruns"] == 20
    assert d["pipeline"]["n_cpus"] == "all"  # default kept by the deep merge
    assert d["simulator"] == {"n_samples": 2000, "delta_t": 0.001, "max_t": 20.0, "smooth_unif": True}
    assert d["training"]["n_samples_per_param"] == 200 and d["training"]["mixture_probabilities"] == [0.8, 0.1, 0.1]
    assert d["estimator"]["type"] == "kde"
    assert d["output"]["folder"] == str(tmp_path / "data/training_data/lan/training_data_n_samples_2000_dt_0.001/angle")
    with pytest.raises(NotImplementedError):
        collect_data_generator_config({**yml, "GENERATOR_APPROACH": "cpn"}, tmp_path)


@pytest.mark.gpu
def test_generate_cli_writes_files(torch_gpu, tmp_path):
    import yaml

    from ssmhip.generate import main

    cfg = tmp_path / "c.yaml"
    cfg.write_text(yaml.safe_dump({"MODEL": "ddm", "PIPELINE": {"N_PARAMETER_SETS": 8},
                                   "SIMULATOR": {"N_SAMPLES": 2000, "DELTA_T": 0.001},
                                   "TRAINING": {"N_SAMPLES_PER_PARAM": 100}}))
    assert main(["--config-path", str(cfg), "--output", str(tmp_path), "-n", "2", "--seed", "3"]) == 0
    files = list((tmp_path / "data").rglob("training_data_*.pickle"))
    assert len(files) == 2
