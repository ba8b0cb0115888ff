"""ssmhip — MI355X-native (gfx950) Euler–Maruyama simulators, KDE likelihood
labelling and a learned likelihood surrogate, with the ssm-simulators API.

Submodules (imported lazily so that host-only code works without a GPU):
  cssm                drop-in ddm_flexbound / race_model / lca / lba_vanilla / lba_angle
  simulator           Simulator(model).simulate(theta, n_samples, ...) and simulator()
  boundary_functions  constant / angle / weibull_cdf
  batched             ragged multi-set launches, hipGraph capture, autotuner
  kde                 LogKDE drop-in and bandwidth_silverman
  estimators          KDELikelihoodEstimator / KDEEstimatorBuilder / ResampleMixtureStrategy
  lan                 on-device LAN training-row generation (simulate -> KDE -> rows)
  surrogate           MLP likelihood surrogate with hand-written BF16-MFMA kernels
"""

__version__ = "0.1.0"

OMISSION_SENTINEL = -999.0


def __getattr__(name):
    import importlib

    if name in {"cssm", "simulator", "boundary_functions", "batched", "kde", "estimators", "lan", "surrogate",
                "config", "em", "autotune"}:
        return importlib.import_module(f"{__name__}.{name}")
    if name in {"Simulator", "simulator_fn"}:
        mod = importlib.import_module(f"{__name__}.simulator")
        return getattr(mod, "Simulator" if name == "Simulator" else "simulator")
    if name == "LogKDE":
        return importlib.import_module(f"{__name__}.kde").LogKDE
    raise AttributeError(name)
