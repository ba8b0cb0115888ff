"""Pins of the MLP oracle (CPU): PyTorch autograd of the same network and finite differences."""

import numpy as np
import pytest
import torch

from oracle import mlp_ref


def _params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for o, n in mlp_ref.LAYERS:
        b = 1 / np.sqrt(n)
        parts += [rng.uniform(-b, b, o * n), rng.uniform(-b, b, o)]
    return np.concatenate(parts).astype(np.float32)


def _torch_net(p):
    net = torch.nn.Sequential(torch.nn.Linear(7, 100), torch.nn.Tanh(), torch.nn.Linear(100, 100), torch.nn.Tanh(),
                              torch.nn.Linear(100, 120), torch.nn.Tanh(), torch.nn.Linear(120, 1)).double()
    lin = [m for m in net if isinstance(m, torch.nn.Linear)]
    with torch.no_grad():
        for (W, b), m in zip(mlp_ref.unflatten(p), lin):
            m.weight.copy_(torch.from_numpy(W.astype(np.float64)))
            m.bias.copy_(torch.from_numpy(b.astype(np.float64)))
    return net, lin


def test_forward_and_backward_match_torch_autograd():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(64, 7)).astype(np.float32)
    gy = rng.normal(size=64)
    y, _ = mlp_ref.forward(p, x)
    dx, dp = mlp_ref.backward(p, x, gy)
    net, lin = _torch_net(p)
    xt = torch.from_numpy(x.astype(np.float64)).requires_grad_(True)
    yt = net(xt)[:, 0]
    assert np.max(np.abs(yt.detach().numpy() - y)) < 1e-12
    (yt * torch.from_numpy(gy)).sum().backward()
    assert np.max(np.abs(xt.grad.numpy() - dx)) < 1e-12
    tg = np.concatenate([np.concatenate([m.weight.grad.numpy().reshape(-1), m.bias.grad.numpy()]) for m in lin])
    assert np.max(np.abs(tg - dp)) < 1e-10


@pytest.mark.parametrize("mode", [1, 2])
def test_rt_transform_forward_and_chain_rule_match_torch_autograd(mode):
    p = _params(5)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.0, 2.0, size=(64, 7)).astype(np.float32)
    x[:8, 5] = x[:8, 3] + np.array([-0.5, -1e-3, 0.0, 5e-5, 2e-4, 1e-2, 0.3, 3.0], np.float32)
    xn = np.concatenate([rng.normal(size=7), rng.uniform(0.5, 2.0, 7), [mode, 3]]).astype(np.float32)
    gy = rng.normal(size=64)
    y, _ = mlp_ref.forward(p, x, xnorm=xn)
    dx, _ = mlp_ref.backward(p, x, gy, xnorm=xn)
    net, _ = _torch_net(p)
    xt = torch.from_numpy(x.astype(np.float64)).requires_grad_(True)
    u = xt[:, 5] - xt[:, 3] if mode == 2 else xt[:, 5]
    feat = torch.log(torch.clamp_min(u, 1e-4))
    xp = torch.cat([xt[:, :5], feat[:, None], xt[:, 6:]], dim=1)
    sh, sc = torch.from_numpy(xn[:7].astype(np.float64)), torch.from_numpy(xn[7:14].astype(np.float64))
    yt = net((xp - sh) * sc)[:, 0]
    assert np.max(np.abs(yt.detach().numpy() - y)) < 1e-5  # float32 standardisation in the oracle
    (yt * torch.from_numpy(gy)).sum().backward()
    assert np.max(np.abs(xt.grad.numpy() - dx)) < 1e-3 * max(1.0, np.abs(dx).max())


def test_backward_matches_central_differences():
    p = _params(3).astype(np.float64)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(5, 7))
    gy = np.ones(5)
    dx, dp = mlp_ref.backward(p.astype(np.float32), x.astype(np.float32), gy)
    p32 = p.astype(np.float32)

    def f(pp, xx):
        return mlp_ref.forward(pp, xx)[0].sum()

    h = 1e-3
    for j in range(7):
        e = np.zeros_like(x)
        e[:, j] = h
        fd = (f(p32, (x + e).astype(np.float32)) - f(p32, (x - e).astype(np.float32))) / (2 * h)
        assert abs(fd - dx[:, j].sum()) < 1e-3 * max(1, abs(fd))
    for k in rng.choice(p.size, 12, replace=False):
        e = np.zeros_like(p32)
        e[k] = h
        fd = (f(p32 + e, x.astype(np.float32)) - f(p32 - e, x.astype(np.float32))) / (2 * h)
        assert abs(fd - dp[k]) < 2e-3 * max(1, abs(fd))


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.0e-3, -2.5], np.float32)
    r = mlp_ref.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0  # halfway between 1.0 and 1.0078125 -> even
    assert r[2] == 1.0078125
    assert np.allclose(r, torch.tensor(v).bfloat16().double().numpy())


def test_sigmoid_carry_rounding_matches_an_independent_torch_restatement():
    """carry="sigmoid": each hidden unit crosses to the next layer as 1 - 2 bf16(u),
    u = sigmoid(-2z); rebuilt here with torch's own bfloat16 rounding and sigmoid."""
    p = _params(7)
    x = np.random.default_rng(8).normal(size=(256, 7)).astype(np.float32)
    y_plain, _ = mlp_ref.forward(p, x)
    assert np.array_equal(mlp_ref.forward(p, x, carry="sigmoid")[0], y_plain)  # no rounding: same network
    y_sig, _ = mlp_ref.forward(p, x, bf16=True, carry="sigmoid")
    y_tanh, _ = mlp_ref.forward(p, x, bf16=True)

    def bf(t):
        return t.float().bfloat16().double()

    s = mlp_ref.TANH_SCALE
    h = bf(torch.from_numpy(x))
    for i, (W, b) in enumerate(mlp_ref.unflatten(p)):
        Wt = torch.from_numpy(W.astype(np.float32))
        Wr = bf(Wt * np.float32(s)) / s if i < 3 else bf(Wt)
        z = h @ Wr.T + torch.from_numpy(b.astype(np.float64))
        if i == 3:
            break
        h = 1.0 - 2.0 * bf(torch.sigmoid(-2.0 * z))
    assert np.max(np.abs(z[:, 0].numpy() - y_sig)) < 1e-9
    # both roundings stay within the bf16 quantisation of the float64 network
    assert np.max(np.abs(y_sig - y_plain)) < 0.05 and np.max(np.abs(y_tanh - y_plain)) < 0.05
    assert np.max(np.abs(y_sig - y_tanh)) > 0  # the rounding point does move the result
    with pytest.raises(ValueError):
        mlp_ref.forward(p, x, carry="relu")
