"""The learned part end to end: train on GPU-produced KDE rows, then show what the
network gains over the method without it (per-evaluation cost against the GPU
KDE) and check its fallback to the KDE outside the training bounds."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(torch_gpu):
    from ssmhip.lan import LanKDELabeller, sample_angle_uniform
    from ssmhip.surrogate import MLPSurrogate

    theta = sample_angle_uniform(512, seed=21)
    lab = LanKDELabeller("angle", theta, n_samples=50_000, seed=3, launch=None)
    rows = lab.run()
    net = MLPSurrogate(seed=1)
    n = rows.shape[0]
    net.fit(rows[: n - 20_000], epochs=30, batch=8192, lr=3e-3, lr_final=1e-4, val_rows=rows[n - 20_000:])
    return net, lab, rows


def test_surrogate_tracks_kde_labels(trained):
    net, _, rows = trained
    # held-out rows: error against the KDE labels (above the floor) well below their spread
    assert net.val_mae is not None
    lab = rows[:, 7]
    spread = float(lab[lab > -65.0].std().item())
    assert net.val_mae < 0.25 * spread, (net.val_mae, spread)


def test_surrogate_is_much_cheaper_than_the_gpu_kde(trained):
    import torch

    net, lab, _ = trained
    bank, pts = lab.bank, lab.points
    x = torch.zeros((pts.rt.shape[0], 7), device="cuda")
    x[:, 5] = pts.rt
    for _ in range(2):
        bank.evaluate(pts)
        net(x)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for _ in range(3):
        bank.evaluate(pts)
    ev[1].record()
    ev[2].record()
    for _ in range(3):
        net(x)
    ev[3].record()
    torch.cuda.synchronize()
    kde_ms = ev[0].elapsed_time(ev[1])
    net_ms = ev[2].elapsed_time(ev[3])
    assert kde_ms / net_ms > 10.0, (kde_ms, net_ms)


def test_fallback_to_kde_outside_training_bounds(trained):
    from ssmhip import _device as D
    from ssmhip.surrogate import SurrogateLikelihood

    net, _, _ = trained
    like = SurrogateLikelihood(net, model="angle", kde_samples=20_000, seed=5, max_val_mae=1e9)
    inside = np.array([[0.5, 1.5, 0.5, 0.3, 0.2, 1.0, 1.0]], np.float32)
    outside = np.array([[0.5, 1.5, 0.5, 0.3, 0.2, 1.0, 1.0], [0.5, 6.0, 0.5, 0.3, 0.2, 2.0, -1.0]], np.float32)
    y_in = like.loglik(D.to_dev(inside)).cpu().numpy()
    assert like.fallback_rows == 0
    assert np.allclose(y_in, net(D.to_dev(inside)).cpu().numpy())
    y_out = like.loglik(D.to_dev(outside)).cpu().numpy()
    assert like.fallback_rows == 1  # a = 6 is outside the trained box
    ref = like.kde_loglik(outside[1:])
    assert np.allclose(y_out[1], ref[0], atol=1e-5)
    assert np.isfinite(y_out).all()
    # above the error threshold every row is labelled by the KDE
    strict = SurrogateLikelihood(net, model="angle", kde_samples=20_000, seed=5, max_val_mae=0.0)
    y_all = strict.loglik(D.to_dev(inside)).cpu().numpy()
    assert strict.fallback_rows == 1
    assert np.allclose(y_all, strict.kde_loglik(inside), atol=1e-5)
