"""The surrogate's targets on a reduced, stated configuration, and gradient-based inference over theta.

Configuration: training rows from this build's LAN pipeline on 16,384 angle sets, theta ~
U(angle bounds), simulated at 1e6 samples each, with 800 KDE-resampled and 200 uniform rows per
set. That is 16.4M rows, 2% of them held out. Training: 100 epochs, Huber loss, Adam, batch 8192,
learning rate 3e-3 -> 5e-5. The bench's own run uses 32,768 sets and 150 epochs.

Targets:
- mean |error| <= 0.05 nats against the analytic DDM log-density at theta = 0;
- inference at least 50x the GPU KDE's evaluations per second on the labelling workload.

Inference over theta: the maximum-likelihood theta of simulated data sets, by gradient ascent
with the hand-written input gradient behind the autograd function. It must recover the true
parameters within the tolerances below, and at theta = 0 agree with the analytic-likelihood
maximum of the same data.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def net(torch_gpu):
    import torch

    from ssmhip.lan import LanKDELabeller, sample_angle_uniform
    from ssmhip.surrogate import MLPSurrogate

    print("[case] surrogate targets: labelling 16384 sets x 1e6", flush=True)
    parts, done = [], 0
    while done < 16384:
        lab = LanKDELabeller("angle", sample_angle_uniform(2048, seed=100 + done), n_samples=1_000_000, n_kde=800,
                             n_unif=200, n_neg=0, seed=200 + done, launch=None)
        parts.append(lab.run())
        del lab
        done += 2048
    rows = torch.cat(parts, 0).contiguous()
    del parts
    n = rows.shape[0]
    perm = torch.randperm(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(0))
    val, train = rows.index_select(0, perm[: n // 50]), rows.index_select(0, perm[n // 50:])
    del rows, perm
    print("[case] surrogate targets: training 100 epochs", flush=True)
    m = MLPSurrogate(seed=0)
    m.fit(train, epochs=100, batch=8192, lr=3e-3, lr_final=5e-5, val_rows=val)
    del train, val
    return m


def _analytic_points(n_sets=2000, per_set=50, seed=5):
    """(x rows [v, a, z, t, 0, rt, choice], analytic log-density): DDM draws at theta = 0."""
    from oracle import ddm_analytic
    from ssmhip import batched

    rng = np.random.default_rng(seed)
    th = rng.uniform([-3.0, 0.3, 0.1, 1e-3], [3.0, 3.0, 0.9, 2.0], (n_sets, 4)).astype(F32)
    sim = batched.BatchedSimulation("ddm", th, per_set, seed=9, launch=None)
    sim.run()
    rts = sim.rts.cpu().numpy().reshape(n_sets, per_set)
    chs = sim.choices.cpu().numpy().reshape(n_sets, per_set)
    xs, ref = [], []
    for i in range(n_sets):
        m = rts[i] > 0
        ll = ddm_analytic.log_density_rt(rts[i][m], chs[i][m], *map(float, th[i]))
        ok = np.isfinite(ll)
        xi = np.zeros((int(ok.sum()), 7), F32)
        xi[:, :4] = th[i]
        xi[:, 5] = rts[i][m][ok]
        xi[:, 6] = chs[i][m][ok]
        xs.append(xi)
        ref.append(ll[ok])
    return np.concatenate(xs), np.concatenate(ref)


def test_mae_against_the_analytic_density_at_theta0(net):
    from ssmhip import _device as D

    x, ref = _analytic_points()
    pred = net(D.to_dev(x)).cpu().numpy()
    mae = float(np.mean(np.abs(pred - ref)))
    print(f"[result] MAE vs analytic at theta = 0: {mae:.4f} nats over {ref.size} points", flush=True)
    assert mae <= 0.05, mae


def test_inference_is_at_least_50x_the_gpu_kde(net):
    import torch

    from ssmhip import batched
    from ssmhip.kde import KDEBank
    from ssmhip.lan import sample_angle_uniform

    # the KDE side: the labelling workload, 1024 angle sets x 1e5 simulations, 1000 points per set
    sim = batched.BatchedSimulation("angle", sample_angle_uniform(1024, seed=1), 100_000, seed=4, smooth_unif=True,
                                    launch=None)
    sim.run()
    bank = KDEBank(sim.rts, sim.choices, sim.job.offsets, [-1, 1])
    pts = bank.sample(800, 200, seed=5)
    rng = np.random.default_rng(6)
    lo, hi = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1]), np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    N = 4_000_000
    x = np.concatenate([rng.uniform(lo, hi, (N, 5)), rng.uniform(0, 20, (N, 1)), rng.choice([-1.0, 1.0], (N, 1))],
                       axis=1).astype(F32)
    dx = torch.from_numpy(x).cuda()
    for _ in range(2):
        bank.evaluate(pts)
        net(dx)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for _ in range(3):
        bank.evaluate(pts)
    ev[1].record()
    ev[2].record()
    for _ in range(3):
        net(dx)
    ev[3].record()
    torch.cuda.synchronize()
    kde_rate = 3 * int(pts.rt.shape[0]) / (ev[0].elapsed_time(ev[1]) / 1e3)
    net_rate = 3 * N / (ev[2].elapsed_time(ev[3]) / 1e3)
    print(f"[result] surrogate {net_rate:.3e} evals/s, GPU KDE {kde_rate:.3e} evals/s: {net_rate / kde_rate:.0f}x",
          flush=True)
    assert net_rate / kde_rate >= 50.0


TRUTHS = [(1.0, 1.5, 0.5, 0.3, 0.2), (-0.8, 1.0, 0.6, 0.5, 0.5), (0.4, 2.0, 0.4, 1.0, 0.0),
          (0.0, 0.6, 0.7, 0.05, 0.3)]
# |theta_hat - theta| per parameter (v, a, z, t, theta) at 3000 trials
TOL = np.array([0.2, 0.2, 0.05, 0.08, 0.12])


@pytest.mark.parametrize("truth", TRUTHS)
def test_gradient_based_theta_inference_recovers_the_parameters(net, truth):
    from ssmhip import batched

    sim = batched.BatchedSimulation("angle", np.array([truth], F32), 3000, seed=32, smooth_unif=True, launch=None)
    sim.run()
    rt = sim.rts.cpu().numpy()
    ch = sim.choices.cpu().numpy().astype(F32)
    ok = rt > 0
    est, ll, starts, lls = net.fit_theta(rt[ok], ch[ok], n_starts=16, steps=300, seed=1)
    err = est - np.array(truth, F32)
    print(f"[result] truth {truth}: estimate {np.round(est, 3)}, error {np.round(err, 3)}", flush=True)
    assert np.all(np.abs(err) <= TOL), (truth, est)
    assert ll >= np.max(lls) - 1e-3 and np.isfinite(ll)


def test_theta_inference_agrees_with_the_analytic_maximum_at_theta0(net):
    """theta = 0 is the plain DDM: the surrogate's maximum-likelihood (v, a, z, t) lies close to
    the maximum of the analytic likelihood (Navarro–Fuss density) of the same trials."""
    from scipy import optimize

    from oracle import ddm_analytic
    from ssmhip import batched

    truth = (0.4, 2.0, 0.4, 1.0, 0.0)
    sim = batched.BatchedSimulation("angle", np.array([truth], F32), 3000, seed=33, smooth_unif=True, launch=None)
    sim.run()
    rt = sim.rts.cpu().numpy()
    ch = sim.choices.cpu().numpy()
    ok = rt > 0
    rt, ch = rt[ok].astype(np.float64), ch[ok]

    def nll(p):
        v, a, z, t = p
        if not (0.3 <= a <= 3.0 and 0.1 <= z <= 0.9 and 0.0 <= t < rt.min()):
            return 1e12
        return -float(np.sum(ddm_analytic.log_density_rt(rt, ch, v, a, z, t)))

    mle = optimize.minimize(nll, x0=np.array(truth[:4]) + [0.1, 0.1, 0.02, -0.02], method="Nelder-Mead",
                            options={"xatol": 1e-4, "fatol": 1e-3, "maxiter": 4000}).x
    est, _, _, _ = net.fit_theta(rt.astype(F32), ch.astype(F32), n_starts=16, steps=300, seed=2,
                                 lower=[-3.0, 0.3, 0.1, 1e-3, -0.1], upper=[3.0, 3.0, 0.9, 2.0, 1.3])
    print(f"[result] analytic MLE {np.round(mle, 3)}, surrogate {np.round(est, 3)}", flush=True)
    assert np.all(np.abs(est[:4] - mle) <= TOL[:4]), (mle, est)
    assert abs(est[4]) <= TOL[4]
