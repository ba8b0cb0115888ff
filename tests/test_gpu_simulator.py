"""The Simulator / cssm drop-ins on the GPU: output contract, preprocessing paths, errors.

Mirrors the reference's API tests (tests/test_simulator.py, tests/test_simulator_class.py):
shapes, dtypes, metadata, same-seed determinism, input formats, deadline handling.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ddm_output_contract(torch_gpu):
    from ssmhip.simulator import Simulator

    out = Simulator("ddm").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.2}, n_samples=1000, random_state=0)
    assert out["rts"].shape == (1000, 1) and out["rts"].dtype == np.float32
    assert out["choices"].shape == (1000, 1) and out["choices"].dtype == np.intc
    md = out["metadata"]
    assert md["model"] == "ddm" and md["simulator"] == "ddm_flexbound" and md["possible_choices"] == [-1, 1]
    assert md["boundary"].shape == (20001,) and np.all(md["boundary"] == np.float32(1.5))
    assert set(np.unique(out["choices"])) <= {-1, 1}
    assert np.all(out["rts"] > 0.2)


def test_input_formats_and_squeeze(torch_gpu):
    import pandas as pd

    from ssmhip.simulator import Simulator

    sim = Simulator("angle")
    theta = {"v": [0.5, -0.5, 1.0], "a": 1.2, "z": 0.5, "t": 0.3, "theta": [0.1, 0.2, 0.3]}
    a = sim.simulate(theta, n_samples=200, random_state=1)
    assert a["rts"].shape == (200, 3, 1)
    arr = np.array([[0.5, 1.2, 0.5, 0.3, 0.1], [-0.5, 1.2, 0.5, 0.3, 0.2], [1.0, 1.2, 0.5, 0.3, 0.3]])
    b = sim.simulate(arr, n_samples=200, random_state=1)
    c = sim.simulate(pd.DataFrame(arr, columns=["v", "a", "z", "t", "theta"]), n_samples=200, random_state=1)
    assert np.array_equal(a["rts"], b["rts"]) and np.array_equal(a["rts"], c["rts"])
    one = sim.simulate(arr, n_samples=1, random_state=2)
    assert one["rts"].shape == (3, 1)  # n_trials > 1 and n_samples == 1 squeezes axis 0


def test_reproducibility_and_seed_validation(torch_gpu):
    from ssmhip.simulator import Simulator

    sim = Simulator("weibull")
    th = {"v": 0.5, "a": 1.5, "z": 0.5, "t": 0.3, "alpha": 2.0, "beta": 3.0}
    r1 = sim.simulate(th, n_samples=500, random_state=42)
    r2 = sim.simulate(th, n_samples=500, random_state=42)
    r3 = sim.simulate(th, n_samples=500, random_state=43)
    assert np.array_equal(r1["rts"], r2["rts"]) and np.array_equal(r1["choices"], r2["choices"])
    assert not np.array_equal(r1["rts"], r3["rts"])
    with pytest.raises(ValueError):
        sim.simulate(th, n_samples=10, random_state=2**31)


def test_deadline_variant_marks_omissions(torch_gpu):
    from ssmhip.simulator import Simulator

    sim = Simulator("ddm_deadline")
    out = sim.simulate({"v": 0.0, "a": 2.5, "z": 0.5, "t": 0.3, "deadline": 1.0}, n_samples=4000, random_state=0)
    rts = out["rts"].reshape(-1)
    assert np.any(rts == -999.0)
    assert np.all((rts == -999.0) | (rts < 1.0))


def test_multi_accumulator_models(torch_gpu):
    from ssmhip.simulator import Simulator

    r = Simulator("race_4").simulate({"v0": 1.0, "v1": 0.5, "v2": 0.5, "v3": 0.2, "a": 2.0, "z0": 0.1, "z1": 0.1,
                                      "z2": 0.1, "z3": 0.1, "t": 0.3}, n_samples=3000, random_state=0)
    assert set(np.unique(r["choices"])) <= {0, 1, 2, 3}
    assert np.mean(r["choices"] == 0) > 0.4
    assert r["metadata"]["possible_choices"] == [0, 1, 2, 3]
    l = Simulator("lca_4").simulate({"v0": 1.0, "v1": 0.5, "v2": 0.5, "v3": 0.2, "a": 2.0, "z0": 0.1, "z1": 0.1,
                                     "z2": 0.1, "z3": 0.1, "g": 0.2, "b": 0.2, "t": 0.3}, n_samples=3000,
                                    random_state=0)
    assert np.mean(l["choices"] == 0) > 0.4
    lba = Simulator("lba3").simulate({"A": 0.3, "b": 0.6, "v0": 0.9, "v1": 0.5, "v2": 0.3}, n_samples=3000,
                                     random_state=0)
    assert lba["rts"].shape == (3000, 1) and np.mean(lba["choices"] == 0) > 0.5
    with pytest.raises(ValueError):
        Simulator("lba3").simulate({"A": 0.7, "b": 0.6, "v0": 0.9, "v1": 0.5, "v2": 0.3}, n_samples=10)


def test_cssm_errors_and_custom_boundary(torch_gpu):
    from ssmhip import cssm
    from ssmhip.simulator import Simulator

    f32 = lambda x: np.array([x], np.float32)  # noqa: E731
    with pytest.raises(ValueError, match="n_threads"):
        cssm.ddm_flexbound(f32(0.5), f32(1.0), f32(0.5), f32(0.3), f32(999.0), f32(1.0), n_threads=300)
    with pytest.raises(ValueError, match="return_option"):
        cssm.ddm_flexbound(f32(0.5), f32(1.0), f32(0.5), f32(0.3), f32(999.0), f32(1.0), return_option="x")

    def linear_collapse(t, a, slope):
        return a - slope * t

    sim = Simulator("ddm", boundary=linear_collapse, boundary_params=["a", "slope"], params=["v", "a", "z", "t",
                                                                                              "slope"])
    out = sim.simulate({"v": 0.5, "a": 1.5, "z": 0.5, "t": 0.3, "slope": 0.2}, n_samples=2000, random_state=0)
    assert out["metadata"]["boundary_fun_type"] == "linear_collapse"
    # a collapsing bound shortens decisions relative to the constant bound
    base = Simulator("ddm").simulate({"v": 0.5, "a": 1.5, "z": 0.5, "t": 0.3}, n_samples=2000, random_state=0)
    assert np.mean(out["rts"]) < np.mean(base["rts"])


def test_graph_replay_matches_eager_launch(torch_gpu):
    import torch

    from ssmhip import batched

    theta, ns = batched.sample_angle_ragged(64, seed=3, lo=100, hi=2000)
    sim = batched.BatchedSimulation("angle", theta, ns, seed=7, launch=None)
    sim.run()
    torch.cuda.synchronize()
    eager = (sim.rts.cpu().numpy().copy(), sim.choices.cpu().numpy().copy())
    for _ in range(3):
        sim.rts.fill_(0)
        sim.replay()
        torch.cuda.synchronize()
        assert np.array_equal(sim.rts.cpu().numpy(), eager[0])
        assert np.array_equal(sim.choices.cpu().numpy(), eager[1])
        torch.zeros(1000, device="cuda").sum()  # unrelated eager work between replays
    sim.close()


def test_graph_of_a_first_time_step_grid_matches_eager_launch(torch_gpu):
    """The walks read their times from a table the library builds once per (device, dt, n_steps).
    A graph captured for a (dt, max_t) no launch of this process has used builds it before the
    capture; its replays equal an eager launch of the same problem afterwards, bit for bit."""
    import torch

    from ssmhip import batched

    theta, ns = batched.sample_angle_ragged(48, seed=21, lo=50, hi=800)
    dt, max_t = 0.000731, 3.37  # a step grid used nowhere else in the tests
    graphed = batched.BatchedSimulation("angle", theta, ns, delta_t=dt, max_t=max_t, seed=5, launch=None)
    print("[case] graph captured before any eager launch of its step grid", flush=True)
    graphed.replay()
    torch.cuda.synchronize()
    got = (graphed.rts.cpu().numpy().copy(), graphed.choices.cpu().numpy().copy())
    graphed.replay()
    torch.cuda.synchronize()
    assert np.array_equal(graphed.rts.cpu().numpy(), got[0])
    graphed.close()
    eager = batched.BatchedSimulation("angle", theta, ns, delta_t=dt, max_t=max_t, seed=5, launch=None)
    eager.run()
    torch.cuda.synchronize()
    assert np.array_equal(eager.rts.cpu().numpy(), got[0])
    assert np.array_equal(eager.choices.cpu().numpy(), got[1])
    assert np.all(got[0][got[0] > 0] <= max_t + 2.0)  # rt = t + non-decision time (< 2 s)


def test_config2_shaped_autotuned_graph_replay_matches_oracle(torch_gpu):
    """1024 ragged angle sets drawn from the batched angle workload's parameter bounds, in one
    launch whose shape the autotuner picks on the batch itself, captured in a HIP graph and
    replayed twice (with unrelated eager work between): every trajectory of every set equals
    the oracle's in replay mode (per-set n_i = floor(10^U(0, log10 40)) keeps the oracle cheap)."""
    import torch

    from oracle import em_numpy as E
    from ssmhip import batched

    K, NMAX, STRIDE = 1024, 40, 257
    theta, ns = batched.sample_angle_ragged(K, seed=12, lo=1, hi=NMAX)
    ns = np.minimum(ns, NMAX)
    offs = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    total = int(offs[-1])
    print(f"[case] config-2-shaped graph replay: {K} sets, {total} trajectories", flush=True)
    rng = np.random.default_rng(13)
    normals = rng.standard_normal((total, STRIDE)).astype(np.float32)
    unif = rng.uniform(size=total).astype(np.float32)
    sim = batched.BatchedSimulation("angle", theta, ns, seed=1, smooth_unif=True, replay_normals=normals,
                                    replay_uniforms=unif, launch="auto", tune_sample=None)
    # the oracle on a uniform NMAX-sample layout; lanes beyond a set's n_i are padding
    pad_n = np.zeros((K * NMAX, STRIDE), np.float32)
    pad_u = np.zeros(K * NMAX, np.float32)
    valid = np.zeros(K * NMAX, bool)
    for k in range(K):
        lo = k * NMAX
        pad_n[lo : lo + ns[k]] = normals[offs[k] : offs[k + 1]]
        pad_u[lo : lo + ns[k]] = unif[offs[k] : offs[k + 1]]
        valid[lo : lo + ns[k]] = True
    v, a, z, t, th = (theta[:, i] for i in range(5))
    r0, c0, s0 = E.ddm_flexbound(v, a, z, t, np.full(K, 999, np.float32), np.ones(K, np.float32), n_samples=NMAX,
                                 delta_t=0.001, boundary="angle", bparams={"a": a, "theta": th}, smooth_unif=True,
                                 normals=pad_n, uniforms=pad_u)
    r0, c0 = r0[valid], c0[valid]
    for rep in range(2):
        sim.rts.fill_(0)
        sim.choices.fill_(9)
        sim.replay()
        torch.cuda.synchronize()
        assert np.array_equal(sim.choices.cpu().numpy().astype(np.int32), c0), rep
        assert np.array_equal(sim.rts.cpu().numpy(), r0), rep
        assert sim.steps() == int(s0[valid].sum())
        torch.zeros(1000, device="cuda").sum()  # unrelated eager work between replays
    sim.close()


def test_full_ddm_and_ornstein_through_the_simulator(torch_gpu):
    from ssmhip.simulator import Simulator

    f = Simulator("full_ddm").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3, "sz": 0.1, "sv": 0.5, "st": 0.1},
                                       n_samples=4000, random_state=0)
    assert f["rts"].shape == (4000, 1) and f["metadata"]["simulator"] == "full_ddm"
    assert f["metadata"]["model"] == "full_ddm"
    assert np.all(f["rts"] > 0.3 - 0.1)  # t_tmp >= t - st
    o = Simulator("ornstein").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "g": 0.5, "t": 0.3}, n_samples=4000,
                                       random_state=0)
    assert o["metadata"]["simulator"] == "ornstein_uhlenbeck" and o["choices"].shape == (4000, 1)
    base = Simulator("ddm").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3}, n_samples=4000, random_state=0)
    # a positive leak pulls the walk back towards v/g: slower and less accurate than the DDM
    assert np.mean(o["rts"]) > np.mean(base["rts"])
    oa = Simulator("ornstein_angle").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "g": 0.5, "t": 0.3, "theta": 0.3},
                                              n_samples=2000, random_state=0)
    assert np.mean(oa["rts"]) < np.mean(o["rts"])  # collapsing bounds shorten decisions


def test_levy_and_racing_diffusion_through_the_simulator(torch_gpu):
    from ssmhip.simulator import Simulator

    lv = Simulator("levy").simulate({"v": 0.5, "a": 1.2, "z": 0.5, "alpha": 1.5, "t": 0.3}, n_samples=4000,
                                    random_state=0)
    assert lv["rts"].shape == (4000, 1) and lv["metadata"]["simulator"] == "levy_flexbound"
    assert lv["metadata"]["model"] == "levy" and set(np.unique(lv["choices"])) <= {-1, 1}
    la = Simulator("levy_angle").simulate({"v": 0.5, "a": 1.2, "z": 0.5, "alpha": 1.5, "t": 0.3, "theta": 0.5},
                                          n_samples=2000, random_state=0)
    assert np.mean(la["rts"]) < np.mean(lv["rts"])  # collapsing bounds shorten decisions
    rd = Simulator("racing_diffusion_3").simulate({"v0": 2.0, "v1": 1.0, "v2": 0.5, "A": 0.5, "b": 1.5, "t": 0.2},
                                                  n_samples=4000, random_state=0)
    assert rd["metadata"]["simulator"] == "rdm_simulator" and rd["metadata"]["possible_choices"] == [0, 1, 2]
    ch = rd["choices"].reshape(-1)
    assert np.mean(ch == 0) > np.mean(ch == 1) > np.mean(ch == 2)  # faster accumulators win more often
    assert np.all(rd["rts"][ch >= 0] >= 0.2)


def test_two_stage_models_through_the_simulator(torch_gpu):
    from ssmhip.simulator import Simulator

    theta = {"vh": 1.0, "vl1": 0.5, "vl2": -0.5, "a": 1.2, "zh": 0.5, "zl1": 0.5, "zl2": 0.5, "t": 0.2}
    p = Simulator("ddm_par2").simulate(theta, n_samples=3000, random_state=0)
    q = Simulator("ddm_seq2").simulate(theta, n_samples=3000, random_state=0)
    assert p["metadata"]["simulator"] == "ddm_flexbound_par2" and q["metadata"]["simulator"] == "ddm_flexbound_seq2"
    assert p["metadata"]["possible_choices"] == [0, 1, 2, 3]
    # the extra arrays keep the simulator's [n_samples, n_trials, 1] shape (only rts and choices are squeezed)
    assert p["rts"].shape == (3000, 1) and p["rts_high"].shape == (3000, 1, 1) and "rts_high" not in q
    assert np.mean(q["rts"]) > np.mean(p["rts"])  # stages in sequence take longer than in parallel
    hi, lo = p["rts_high"][:, 0], p["rts_low"][:, 0]
    assert np.all(p["rts"] >= hi - 1e-3) and np.all(p["rts"] >= lo - 1e-3)
    for name in ("ddm_par2_no_bias", "ddm_seq2_angle_no_bias", "ddm_par2_weibull_no_bias",
                 "ddm_seq2_conflict_gamma_no_bias"):
        cfg = Simulator(name).config
        th = dict(zip(cfg["params"], cfg["default_params"]))
        out = Simulator(name).simulate(th, n_samples=500, random_state=1)
        assert set(np.unique(out["choices"])) <= {0, 1, 2, 3}, name


def test_legacy_ddm_through_the_simulator(torch_gpu):
    from ssmhip.simulator import Simulator

    out = Simulator("ddm_legacy").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3}, n_samples=4000, random_state=0)
    assert out["metadata"]["simulator"] == "ddm" and "boundary" not in out["metadata"]
    assert set(np.unique(out["choices"])) <= {-1, 1} and np.all(out["rts"] > 0.3)
    base = Simulator("ddm").simulate({"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.3}, n_samples=4000, random_state=0)
    assert np.mean(out["rts"]) < np.mean(base["rts"])  # bounds a apart instead of 2a


# One model per simulator family, with two parameter rows that differ only in boundary
# parameters (the reference's tests/test_metadata_boundary_trial0.py uses the same models).
BOUNDARY_ROWS = {
    "angle": {"theta": (0.0, 1.2)},
    "weibull_cdf": {"alpha": (3.0, 1.0), "beta": (3.0, 1.0)},
    "ddm_seq2_angle_no_bias": {"theta": (0.0, 1.0)},
    "race_no_bias_angle_4": {"theta": (0.0, 1.0)},
    "levy_angle": {"theta": (0.0, 1.0)},
    "ornstein_angle": {"theta": (0.0, 1.0)},
    "ddm_par2_angle_no_bias": {"theta": (0.0, 1.0)},
}


@pytest.mark.parametrize("model", list(BOUNDARY_ROWS))
def test_full_metadata_boundary_describes_trial_0(torch_gpu, model):
    """metadata['boundary'] is the boundary of the first parameter row, whatever the other
    rows hold and whatever n_threads says, and equals the boundary function on the time grid."""
    from ssmhip import config
    from ssmhip.boundary_functions import time_grid
    from ssmhip.simulator import simulator

    print(f"[case] metadata boundary {model}", flush=True)
    cfg = config.get(model)
    row0 = dict(zip(cfg["params"], cfg["default_params"]))
    row1 = dict(row0)
    for p, (x0, x1) in BOUNDARY_ROWS[model].items():
        row0[p], row1[p] = x0, x1

    def bnd(rows, n_threads=1):
        theta = {k: np.array([r[k] for r in rows], dtype=np.float32) for k in rows[0]}
        out = simulator(theta, model=model, n_samples=2, random_state=42, n_threads=n_threads)
        return np.asarray(out["metadata"]["boundary"])

    b0, b1 = bnd([row0]), bnd([row1])
    assert not np.allclose(b0, b1)
    np.testing.assert_allclose(bnd([row0, row1]), b0, rtol=1e-6)
    np.testing.assert_allclose(bnd([row1, row0]), b1, rtol=1e-6)
    np.testing.assert_allclose(bnd([row0, row1], n_threads=4), b0, rtol=1e-6)
    names = cfg["boundary_params"]
    expect = cfg["boundary"](t=time_grid(20.0, 0.001), **{k: np.float32(row0[k]) for k in names})
    np.testing.assert_allclose(b0, np.asarray(expect, np.float32), rtol=1e-6, atol=1e-7)


def test_race_and_lca_variants_through_the_simulator(torch_gpu):
    """Every race/LCA variant configuration runs at its default parameters; an angle
    bound that collapses faster ends the race sooner."""
    from ssmhip import config
    from ssmhip.simulator import Simulator

    for name in [m for m in config.list_models() if m.startswith(("race_", "lca_"))]:
        print(f"[case] {name}", flush=True)
        cfg = config.get(name)
        theta = dict(zip(cfg["params"], cfg["default_params"]))
        theta.update({f"v{i}": 1.0 for i in range(cfg["n_particles"])})
        out = Simulator(name).simulate(theta, n_samples=500, random_state=1)
        assert out["rts"].shape == (500, 1) and out["choices"].dtype == np.intc
        assert set(np.unique(out["choices"])) <= set(cfg["choices"])
        assert np.all(out["rts"] > 0) and out["metadata"]["possible_choices"] == cfg["choices"]
    base = {"v0": 1.0, "v1": 0.8, "v2": 0.6, "v3": 0.4, "a": 2.0, "z": 0.2, "t": 0.1}
    slow = Simulator("race_no_bias_angle_4").simulate({**base, "theta": 0.0}, n_samples=4000, random_state=2)
    fast = Simulator("race_no_bias_angle_4").simulate({**base, "theta": 1.0}, n_samples=4000, random_state=2)
    assert fast["rts"].mean() < slow["rts"].mean() - 0.2
    const = Simulator("race_no_bias_4").simulate(base, n_samples=4000, random_state=2)
    assert abs(const["rts"].mean() - slow["rts"].mean()) < 0.05


def test_zero_samples_give_empty_outputs(torch_gpu):
    """n_samples = 0 launches nothing and returns empty arrays of the reference's shapes."""
    from ssmhip.simulator import Simulator

    for name, theta in (("angle", {"v": 1.0, "a": 1.5, "z": 0.5, "t": 0.2, "theta": 0.3}),
                        ("race_no_bias_3", {"v0": 1.0, "v1": 0.5, "v2": 0.2, "a": 2.0, "z": 0.3, "t": 0.1}),
                        ("lba2", {"A": 0.3, "b": 0.6, "v0": 0.9, "v1": 0.5}),
                        ("ddm_seq2_no_bias", {"vh": 0.5, "vl1": 0.5, "vl2": -0.5, "a": 1.0, "t": 0.2})):
        print(f"[case] zero samples {name}", flush=True)
        out = Simulator(name).simulate(theta, n_samples=0, random_state=3)
        assert out["rts"].shape[0] == 0 and out["rts"].dtype == np.float32
        assert out["choices"].shape == out["rts"].shape and out["choices"].dtype == np.intc
