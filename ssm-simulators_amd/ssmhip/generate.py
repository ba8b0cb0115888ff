"""`python -m ssmhip.generate`: LAN training-data files from a YAML configuration, on the GPU.

Reads the YAML layout of ssms/cli/config_data_generation.yaml (MODEL, GENERATOR_APPROACH,
PIPELINE, SIMULATOR, TRAINING, ESTIMATOR; keys case-insensitive) and builds the nested
generator configuration as collect_data_generator_config does (ssms/cli/generate.py:150-215):
the YAML's PIPELINE section forwarded whole, SIMULATOR n_samples / delta_t, TRAINING
n_samples_per_param, ESTIMATOR type, and the output folder
<output>/data/training_data/<approach>/training_data_n_samples_<n>_dt_<dt>/<model>
(generate.py:87-97). Each file is one TrainingDataGenerator.generate_data_training(save=True)
call of ssmhip.training_data. MLflow logging and the worker-process options of the reference
CLI have no counterpart here (the batch runs on one GPU).
"""

from __future__ import annotations

import argparse
import logging
import sys
from pathlib import Path

import yaml

from . import config as model_configs
from .training_data import TrainingDataGenerator, get_lan_kde_config


def _lower_keys(d):
    return {str(k).lower(): _lower_keys(v) if isinstance(v, dict) else v for k, v in d.items()}


def data_folder(base_path, approach: str, n_samples, delta_t, model: str) -> Path:
    return (Path(base_path) / "data/training_data" / approach / f"training_data_n_samples_{n_samples}_dt_{delta_t}"
            / model)


def collect_data_generator_config(yaml_config: dict, base_path) -> dict:
    """{"model_config", "data_config"} from a parsed YAML (generate.py:150-215 + make_data_generator_configs)."""
    bc = _lower_keys(yaml_config)
    model = bc.get("model", "ddm")
    approach = bc.get("generator_approach", "lan")
    if approach != "lan":
        raise NotImplementedError(f"generator approach {approach!r}: only 'lan' runs on the device")
    data = get_lan_kde_config(model)
    sim = bc.get("simulator", {})
    nested = {"output": {"folder": str(data_folder(base_path, approach, sim.get("n_samples"), sim.get("delta_t"),
                                                   model))},
              "model": model, "cpn_only": False}
    if "pipeline" in bc:
        nested["pipeline"] = dict(bc["pipeline"])
    if sim:
        nested["simulator"] = {k: sim[k] for k in ("n_samples", "delta_t") if k in sim}
    if "training" in bc:
        nested["training"] = {"n_samples_per_param": bc["training"]["n_samples_per_param"]}
    if "estimator" in bc:
        nested["estimator"] = {"type": str(bc["estimator"].get("type", "kde")).lower()}
    for section, values in nested.items():  # deep merge into the defaults
        if section in data and isinstance(values, dict):
            data[section].update(values)
        else:
            data[section] = values
    data["generator_approach"] = approach
    return {"model_config": model_configs.get(model), "data_config": data}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m ssmhip.generate", description=__doc__.splitlines()[0])
    ap.add_argument("--config-path", type=Path, required=True, help="YAML configuration file")
    ap.add_argument("--output", type=Path, required=True, help="base output directory")
    ap.add_argument("--n-files", "-n", type=int, default=1, help="number of files to generate")
    ap.add_argument("--dry-run", action="store_true", help="generate one batch without saving")
    ap.add_argument("--seed", type=int, default=None, help="seed of the first file (default: random)")
    ap.add_argument("--log-level", default="INFO")
    args = ap.parse_args(argv)
    logging.basicConfig(level=args.log_level.upper(), format="%(asctime)s - %(levelname)s - %(message)s")
    log = logging.getLogger("ssmhip.generate")
    if args.n_files < 1:
        log.error("--n-files must be >= 1")
        return 2
    with open(args.config_path, "rb") as f:
        cfg = collect_data_generator_config(yaml.safe_load(f), args.output)
    gen = TrainingDataGenerator(cfg["data_config"], cfg["model_config"], seed=args.seed)
    folder = Path(cfg["data_config"]["output"]["folder"])
    if args.dry_run:
        gen.generate_data_training(save=False)
        log.info("dry run: pipeline ok; would write %d file(s) to %s", args.n_files, folder)
        return 0
    before = set(folder.rglob("*.pickle")) if folder.exists() else set()
    for _ in range(args.n_files):
        gen.generate_data_training(save=True)
    new = sorted(set(folder.rglob("*.pickle")) - before)
    log.info("wrote %d file(s) to %s", len(new), folder)
    for p in new:
        print(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
