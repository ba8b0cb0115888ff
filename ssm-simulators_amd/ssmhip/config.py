"""Model configurations for the models whose simulators run on the GPU.

Same dictionary contract as the reference's registry entries
(ssms/config/_modelconfig/{ddm,angle,weibull,race,lca,lba}.py): name, params,
param_bounds, boundary_name, boundary, boundary_params, n_params,
default_params, nchoices, choices, n_particles, simulator, and the
simulation-time parameter transforms (ColumnStackParameters / ExpandDimension
/ renames of ssms/transforms/simulation/common.py) expressed as callables
``f(theta, config, n_trials) -> theta``.
"""

from __future__ import annotations

from copy import deepcopy

import numpy as np

from . import boundary_functions as bf
from . import cssm

DEADLINE_PARAM = {"name": "deadline", "default_bounds": (0.001, 10.0), "default_value": 10.0}


def column_stack(sources, target):
    def f(theta, cfg, n):
        theta[target] = np.column_stack([theta[s] for s in sources]).astype(np.float32)
        return theta

    return f


def expand_dims(names):
    def f(theta, cfg, n):
        for p in names:
            if p in theta:
                theta[p] = np.expand_dims(theta[p], axis=1)
        return theta

    return f


def _lba_transforms(nact):
    def f(theta, cfg, n):
        theta["nact"] = nact
        theta["v"] = np.column_stack([theta[f"v{i}"] for i in range(nact)]).astype(np.float32)
        theta["z"] = np.expand_dims(theta.pop("A"), axis=1)
        theta["a"] = np.expand_dims(theta.pop("b"), axis=1)
        theta["t"] = np.zeros(n).astype(np.float32)
        return theta

    return [f]


def _lba_angle_3_transforms():
    def f(theta, cfg, n):
        theta["v"] = np.column_stack([theta[f"v{i}"] for i in range(3)]).astype(np.float32)
        for p in ("a", "z", "theta"):
            theta[p] = np.expand_dims(theta[p], axis=1)
        theta["t"] = np.zeros(n).astype(np.float32)
        return theta

    return [f]


def _ddm_like(name, params, bounds, defaults, bname, simulator=None):
    fun, bparams, _ = bf.BOUNDARY_REGISTRY[bname]
    return {
        "name": name, "params": params, "param_bounds": bounds, "boundary_name": bname, "boundary": fun,
        "boundary_params": [] if bname == "constant" else None, "n_params": len(params),
        "default_params": defaults, "nchoices": 2, "choices": [-1, 1], "n_particles": 1,
        "simulator": simulator or cssm.ddm_flexbound, "simulation_transforms": [],
    }


def _race_z(n, zs, variant):
    """z handling of the race/LCA variants: one z per accumulator (ColumnStackParameters),
    one shared z duplicated n times (no_bias), or zeros (no_z; SetZeroArray)."""
    def f(theta, cfg, n_trials):
        if variant == "":
            theta["z"] = np.column_stack([theta[s] for s in zs]).astype(np.float32)
        elif variant == "no_bias":
            theta["z"] = np.column_stack([theta["z"]] * n).astype(np.float32)
        else:
            theta["z"] = np.zeros((n_trials, n), dtype=np.float32)
        return theta

    return f


def _race(n, lca=False, variant="", angle=False):
    """race_{n} / lca_{n} and their no_bias / no_z / *_angle variants (race.py:14-480,
    lca.py:15-330): same simulator, z from one column per accumulator, one shared
    column or zeros, and the angle boundary (theta in [-0.1, 1.45])."""
    vs = [f"v{i}" for i in range(n)]
    zs = {"": [f"z{i}" for i in range(n)], "no_bias": ["z"], "no_z": []}[variant]
    extra = ["g", "b"] if lca else []
    params = vs + ["a"] + zs + extra + ["t"] + (["theta"] if angle else [])
    # lca.py gives theta a lower bound of -1.0 in its three-accumulator angle variants
    theta_lo = -1.0 if (lca and angle and n == 3) else -0.1
    lo = [0.0] * n + [1.0] + [0.0] * len(zs) + ([-1.0, -1.0] if lca else []) + [0.0] + ([theta_lo] if angle else [])
    hi = [2.5] * n + [3.0] + [0.9] * len(zs) + ([1.0, 1.0] if lca else []) + [2.0] + ([1.45] if angle else [])
    dflt = [0.0] * n + [2.0] + [0.5] * len(zs) + ([0.0, 0.0] if lca else []) + [1e-3] + ([0.0] if angle else [])
    tag = (variant + "_" if variant else "") + ("angle_" if angle else "")
    bname = "angle" if angle else "constant"
    return {
        "name": f"{'lca' if lca else 'race'}_{tag}{n}", "params": params, "param_bounds": [lo, hi],
        "boundary_name": bname, "boundary": bf.angle if angle else bf.constant, "n_params": len(params),
        "default_params": dflt, "nchoices": n, "choices": list(range(n)), "n_particles": n,
        "simulator": cssm.lca if lca else cssm.race_model,
        "simulation_transforms": [_race_z(n, zs, variant), column_stack(vs, "v"),
                                  expand_dims(["t", "a"] + (["g", "b"] if lca else []))],
    }


def _lba(n, bounds, defaults):
    params = ["A", "b"] + [f"v{i}" for i in range(n)]
    return {
        "name": f"lba{n}", "params": params, "param_bounds": bounds, "boundary_name": "constant",
        "boundary": bf.constant, "n_params": len(params), "default_params": defaults, "nchoices": n,
        "choices": list(range(n)), "n_particles": n, "simulator": cssm.lba_vanilla,
        "simulation_transforms": _lba_transforms(n),
    }


def _z_half():
    def f(theta, cfg, n):
        for p in ("zh", "zl1", "zl2"):
            theta[p] = np.tile(np.array([0.5], dtype=np.float32), n)
        return theta

    return [f]


def _two_stage(kind, variant):
    """ddm_par2.py / ddm_seq2.py: the full model and the no-bias variants (z's fixed at 1/2)
    with constant, conflict_gamma, angle and weibull_cdf bounds."""
    v3 = ["vh", "vl1", "vl2"]
    lo3, hi3 = [-4.0] * 3, [4.0] * 3
    table = {
        "": (v3 + ["a", "zh", "zl1", "zl2", "t"], [lo3 + [0.3, 0.2, 0.2, 0.2, 0.0], hi3 + [2.5, 0.8, 0.8, 0.8, 2.0]],
             [0.0, 0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 1.0], "constant"),
        "_no_bias": (v3 + ["a", "t"], [lo3 + [0.3, 0.0], hi3 + [2.5, 2.0]], [0.0, 0.0, 0.0, 1.0, 1.0], "constant"),
        "_conflict_gamma_no_bias": (
            v3 + ["t", "a", "theta", "scale", "alphaGamma", "scaleGamma"],
            [lo3 + [0.0, 0.3, 0.0, 0.0, 1.1, 0.5], hi3 + [2.0, 2.5, 0.5, 5.0, 5.0, 5.0]],
            [0.0, 0.0, 0.0, 1.0, 1.0, 0.5, 1.0, 2, 2], "conflict_gamma"),
        "_angle_no_bias": (v3 + ["a", "t", "theta"], [lo3 + [0.3, 0.0, -0.1], hi3 + [2.5, 2.0, 1.0]],
                           [0.0, 0.0, 0.0, 1.0, 1.0, 0.0], "angle"),
        "_weibull_no_bias": (v3 + ["a", "t", "alpha", "beta"], [lo3 + [0.3, 0.0, 0.31, 0.31], hi3 + [2.5, 2.0, 4.99, 6.99]],
                             [0.0, 0.0, 0.0, 1.0, 1.0] + ([0.0, 3.0] if kind == "par2" else [2.5, 3.5]), "weibull_cdf"),
    }
    params, bounds, dflt, bname = table[variant]
    fun, bnames, _ = bf.BOUNDARY_REGISTRY[bname]
    return {
        "name": f"ddm_{kind}{variant}", "params": params, "param_bounds": bounds, "boundary_name": bname,
        "boundary": fun, "boundary_params": list(bnames), "n_params": len(params), "default_params": dflt,
        "nchoices": 4, "choices": [0, 1, 2, 3], "n_particles": 1,
        "simulator": cssm.ddm_flexbound_par2 if kind == "par2" else cssm.ddm_flexbound_seq2,
        "simulation_transforms": [] if variant == "" else _z_half(),
    }


def _registry() -> dict:
    reg = {
        # ddm.py:7-26, angle.py:7-25, weibull.py:7-28
        "ddm": _ddm_like("ddm", ["v", "a", "z", "t"], [[-3.0, 0.3, 0.1, 0.0], [3.0, 2.5, 0.9, 2.0]],
                         [0.0, 1.0, 0.5, 1e-3], "constant"),
        "angle": _ddm_like("angle", ["v", "a", "z", "t", "theta"],
                           [[-3.0, 0.3, 0.1, 1e-3, -0.1], [3.0, 3.0, 0.9, 2.0, 1.3]], [0.0, 1.0, 0.5, 1e-3, 0.0],
                           "angle"),
        "weibull": _ddm_like("weibull", ["v", "a", "z", "t", "alpha", "beta"],
                             [[-2.5, 0.3, 0.2, 1e-3, 0.31, 0.31], [2.5, 2.5, 0.8, 2.0, 4.99, 6.99]],
                             [0.0, 1.0, 0.5, 1e-3, 3.0, 3.0], "weibull_cdf"),
        # ddm.py:30-48 (the [0, a]-bound legacy simulator)
        "ddm_legacy": _ddm_like("ddm_legacy", ["v", "a", "z", "t"], [[-3.0, 0.3, 0.1, 0.0], [3.0, 2.5, 0.9, 2.0]],
                                [0.0, 1.0, 0.5, 1e-3], "constant", cssm.ddm),
        # full_ddm.py:11-33, ornstein.py:7-48
        "full_ddm": _ddm_like("full_ddm", ["v", "a", "z", "t", "sz", "sv", "st"],
                              [[-3.0, 0.3, 0.3, 0.25, 1e-3, 1e-3, 1e-3], [3.0, 2.5, 0.7, 2.25, 0.2, 2.0, 0.25]],
                              [0.0, 1.0, 0.5, 0.25, 1e-3, 1e-3, 1e-3], "constant", cssm.full_ddm),
        "ornstein": _ddm_like("ornstein", ["v", "a", "z", "g", "t"], [[-2.0, 0.3, 0.1, -1.0, 1e-3], [2.0, 3.0, 0.9, 1.0, 2]],
                              [0.0, 1.0, 0.5, 0.0, 1e-3], "constant", cssm.ornstein_uhlenbeck),
        "ornstein_angle": _ddm_like("ornstein_angle", ["v", "a", "z", "g", "t", "theta"],
                                    [[-2.0, 0.3, 0.1, -1.0, 1e-3, -0.1], [2.0, 3.0, 0.9, 1.0, 2, 1.3]],
                                    [0.0, 1.0, 0.5, 0.0, 1e-3, 0.1], "angle", cssm.ornstein_uhlenbeck),
        # levy.py:7-48
        "levy": _ddm_like("levy", ["v", "a", "z", "alpha", "t"], [[-3.0, 0.3, 0.1, 1.0, 1e-3], [3.0, 3.0, 0.9, 2.0, 2]],
                          [0.0, 1.0, 0.5, 1.5, 0.1], "constant", cssm.levy_flexbound),
        "levy_angle": _ddm_like("levy_angle", ["v", "a", "z", "alpha", "t", "theta"],
                                [[-3.0, 0.3, 0.1, 1.0, 1e-3, -0.1], [3.0, 3.0, 0.9, 2.0, 2, 1.3]],
                                [0.0, 1.0, 0.5, 1.5, 0.1, 0.01], "angle", cssm.levy_flexbound),
        # racing_diffusion.py:11-36
        "racing_diffusion_3": {
            "name": "racing_diffusion_3", "params": ["v0", "v1", "v2", "A", "b", "t"],
            "param_bounds": [[0.0, 0.0, 0.0, 0.0, 0.5, 0.0], [2.5, 2.5, 2.5, 1.0, 3.0, 2.0]],
            "boundary_name": "constant", "boundary": bf.constant, "boundary_params": [], "n_params": 6,
            "default_params": [1.0, 1.0, 1.0, 0.5, 1.5, 1e-3], "nchoices": 3, "choices": [0, 1, 2],
            "n_particles": 3, "simulator": cssm.racing_diffusion_model,
            "simulation_transforms": [column_stack(["v0", "v1", "v2"], "v"), expand_dims(["t", "A", "b"])],
        },
        # lba.py:94-222
        "lba2": _lba(2, [[0.0, 0.0, 0.0, 0.1], [1.0, 1.0, 1.0, 1.1]], [0.3, 0.5, 0.5, 0.5]),
        "lba3": _lba(3, [[0.0, 0.0, 0.0, 0.1, 0.1], [1.0, 1.0, 1.0, 1.1, 0.50]], [0.3, 0.5, 0.25, 0.5, 0.25]),
        "lba4": _lba(4, [[0.0, 0.0, 0.0, 0.1, 0.1, 0.1], [1.0, 1.0, 1.0, 1.1, 0.50, 0.50]],
                     [0.3, 0.5, 0.25, 0.25, 0.25, 0.25]),
        "lba_angle_3": {
            "name": "lba_angle_3", "params": ["v0", "v1", "v2", "a", "z", "theta"],
            "param_bounds": [[0.0, 0.0, 0.0, 0.1, 0.0, 0], [6.0, 6.0, 6.0, 1.1, 0.5, 1.3]],
            "boundary_name": "constant", "boundary": bf.constant, "n_params": 6,
            "default_params": [0.5, 0.3, 0.2, 0.5, 0.2, 0.0], "nchoices": 3, "choices": [0, 1, 2],
            "n_particles": 3, "simulator": cssm.lba_angle, "simulation_transforms": _lba_angle_3_transforms(),
        },
    }
    # race.py:14-480 (2-4 accumulators), lca.py:15-330 (3-4): five variants each
    for variant, angle in (("", False), ("no_bias", False), ("no_z", False), ("no_bias", True), ("no_z", True)):
        for n in (2, 3, 4):
            cfg = _race(n, False, variant, angle)
            reg[cfg["name"]] = cfg
        for n in (3, 4):
            cfg = _race(n, True, variant, angle)
            reg[cfg["name"]] = cfg
    # _modelconfig/__init__.py registers the weibull and full-DDM configurations under
    # a second name as well
    reg["weibull_cdf"] = reg["weibull"]
    reg["full_ddm2"] = reg["full_ddm"]
    # ddm_par2.py:10-185, ddm_seq2.py:11-190
    for kind in ("par2", "seq2"):
        for variant in ("", "_no_bias", "_conflict_gamma_no_bias", "_angle_no_bias", "_weibull_no_bias"):
            cfg = _two_stage(kind, variant)
            reg[cfg["name"]] = cfg
    for cfg in reg.values():
        if cfg.get("boundary_params") is None:
            cfg["boundary_params"] = list(bf.BOUNDARY_REGISTRY[cfg["boundary_name"]][1])
    return reg


_REGISTRY = _registry()


def list_models() -> list[str]:
    return sorted(_REGISTRY)


def with_deadline(config: dict) -> dict:
    """ModelConfigBuilder.with_deadline (ssms/config/model_config_builder.py:471-560)."""
    if "deadline" in config.get("params", []):
        return deepcopy(config)
    cfg = deepcopy(config)
    cfg["params"] = cfg["params"] + ["deadline"]
    cfg["n_params"] = cfg["n_params"] + 1
    if not cfg["name"].endswith("_deadline"):
        cfg["name"] = cfg["name"] + "_deadline"
    lo, hi = cfg["param_bounds"]
    cfg["param_bounds"] = [lo + [DEADLINE_PARAM["default_bounds"][0]], hi + [DEADLINE_PARAM["default_bounds"][1]]]
    cfg["default_params"] = cfg["default_params"] + [DEADLINE_PARAM["default_value"]]
    cfg["deadline"] = True
    return cfg


def get(model_name: str, **overrides) -> dict:
    """ModelConfigBuilder.from_model: registry lookup + '_deadline' variant + overrides."""
    base = model_name.replace("_deadline", "")
    if base not in _REGISTRY:
        raise ValueError(f"Unknown model '{base}'. Available models: {list_models()}")
    cfg = deepcopy(_REGISTRY[base])
    cfg.update(overrides)
    if "_deadline" in model_name:
        cfg = with_deadline(cfg)
    return cfg
