"""GPU checks of the BF16-MFMA surrogate kernels.

Forward / backward against the NumPy oracle with the same bf16 operand
rounding (tolerance set by float32 vs float64 accumulation order and the
hardware exp/rcp inside tanh), against PyTorch's autograd of the same network
in float32 (tolerance set by bf16 operands, ~1e-2 relative), and the input
gradient against central finite differences of the float64 network.
"""

import math

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    lo = np.array([-3.0, 0.3, 0.1, 1e-3, -0.1])
    hi = np.array([3.0, 3.0, 0.9, 2.0, 1.3])
    th = rng.uniform(lo, hi, (n, 5))
    rt = rng.uniform(0.0, 20.0, (n, 1))
    ch = rng.choice([-1.0, 1.0], (n, 1))
    return np.concatenate([th, rt, ch], axis=1).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


def test_mfma_lane_layout(torch_gpu):
    from ssmhip import _device as D
    from ssmhip import _lib as L

    torch = torch_gpu
    rng = np.random.default_rng(0)
    A = mlp_ref.bf16_round(rng.normal(size=(16, 32)).astype(np.float32))
    B = mlp_ref.bf16_round(rng.normal(size=(32, 16)).astype(np.float32))
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for j in range(8):
            a[l, j] = A[l % 16, 8 * (l // 16) + j]
            b[l, j] = B[8 * (l // 16) + j, l % 16]
    da = D.to_dev(a).to(torch.bfloat16)
    db = D.to_dev(b).to(torch.bfloat16)
    dd = D.empty((64, 4), torch.float32)
    L.check(L.load().ssm_mfma_probe(D.ptr(da), D.ptr(db), D.ptr(dd), D.stream_ptr()))
    d = dd.cpu().numpy()
    ref = A @ B
    for l in range(64):
        for i in range(4):
            assert abs(d[l, i] - ref[4 * (l // 16) + i, l % 16]) < 1e-3


def test_mfma32_lane_layout(torch_gpu):
    """The inference kernel's v_mfma_f32_32x32x16_bf16: A[m = l % 32][k = 8 (l // 32) + i],
    B[k = 8 (l // 32) + i][n = l % 32], result register r = D[8 (r // 4) + 4 (l // 32) + r % 4][l % 32]."""
    from ssmhip import _device as D
    from ssmhip import _lib as L

    torch = torch_gpu
    rng = np.random.default_rng(1)
    A = mlp_ref.bf16_round(rng.normal(size=(32, 16)).astype(np.float32))
    B = mlp_ref.bf16_round(rng.normal(size=(16, 32)).astype(np.float32))
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for j in range(8):
            a[l, j] = A[l % 32, 8 * (l // 32) + j]
            b[l, j] = B[8 * (l // 32) + j, l % 32]
    da = D.to_dev(a).to(torch.bfloat16)
    db = D.to_dev(b).to(torch.bfloat16)
    dd = D.empty((64, 16), torch.float32)
    L.check(L.load().ssm_mfma32_probe(D.ptr(da), D.ptr(db), D.ptr(dd), D.stream_ptr()))
    d = dd.cpu().numpy()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    for l in range(64):
        for r in range(16):
            assert abs(d[l, r] - ref[8 * (r // 4) + 4 * (l // 32) + r % 4, l % 32]) < 1e-3


def test_forward_matches_oracle_and_torch(torch_gpu):
    from ssmhip import _device as D
    from ssmhip.surrogate import MLPSurrogate

    torch = torch_gpu
    net = MLPSurrogate(seed=1)
    x = _inputs(4099, seed=1)  # not a multiple of 32: exercises the row tail
    y = net(D.to_dev(x)).cpu().numpy()
    p = net.params.cpu().numpy()
    yb, _ = mlp_ref.forward(p, x, bf16=True, carry="sigmoid")  # the inference kernel carries u = (1 - tanh) / 2
    assert np.max(np.abs(y - yb)) < 5e-3
    # and it rounds u, not tanh: the tanh-rounding model is much further off
    yh, _ = mlp_ref.forward(p, x, bf16=True)
    assert np.mean(np.abs(y - yb)) < 0.5 * np.mean(np.abs(y - yh))
    yt = net.torch_reference().cuda()(torch.from_numpy(x).cuda())[:, 0].detach().cpu().numpy()
    # against the unrounded network the error is the bf16 rounding of u: 2^-9 absolute where
    # tanh is near 0 (random initial weights keep most units there), as the rounding model says
    y64, _ = mlp_ref.forward(p, x)
    assert _rel(yb, y64) < 4e-2 and abs(_rel(y, yt) - _rel(yb, y64)) < 5e-3


def test_autograd_backward_matches_oracle_and_torch(torch_gpu):
    from ssmhip import _device as D
    from ssmhip.surrogate import MLPSurrogate, param_slices

    torch = torch_gpu
    net = MLPSurrogate(seed=2)
    x_np = _inputs(1000, seed=2)
    x = D.to_dev(x_np).requires_grad_(True)
    params = net.params.clone().requires_grad_(True)
    from ssmhip.surrogate import mlp_apply

    y = mlp_apply(x, params, net.kern)
    gy = np.random.default_rng(3).normal(size=1000).astype(np.float32)
    (y * D.to_dev(gy)).sum().backward()
    dx = x.grad.cpu().numpy()
    dp = params.grad.cpu().numpy()
    p = net.params.cpu().numpy()
    dx_b, dp_b = mlp_ref.backward(p, x_np, gy.astype(np.float64), bf16=True)
    assert _rel(dx, dx_b) < 1e-2
    assert _rel(dp, dp_b) < 1e-2
    # PyTorch autograd of the same network, float32, no bf16 rounding
    ref = net.torch_reference().cuda()
    xt = torch.from_numpy(x_np).cuda().requires_grad_(True)
    (ref(xt)[:, 0] * torch.from_numpy(gy).cuda()).sum().backward()
    assert _rel(dx, xt.grad.cpu().numpy()) < 3e-2
    lin = [m for m in ref if isinstance(m, torch.nn.Linear)]
    tg = np.concatenate([np.concatenate([m.weight.grad.cpu().numpy().reshape(-1), m.bias.grad.cpu().numpy()])
                         for m in lin])
    assert _rel(dp, tg) < 3e-2
    for name, sl, _ in param_slices():
        assert _rel(dp[sl], tg[sl]) < 6e-2, name


@pytest.mark.parametrize("rt_mode", ["log", "log_minus_t"])
def test_standardised_inputs_forward_and_gradients(torch_gpu, rt_mode):
    from ssmhip import _device as D
    from ssmhip.surrogate import MLPSurrogate

    net = MLPSurrogate(seed=6)
    x_np = _inputs(2000, seed=6)
    x_np[:7, 5] = [0.0, -0.5, 5e-5, 1.5e-4, 2e-4, 0.01, 19.0]  # rt below, near and above the log floor
    x_np[7:12, 5] = x_np[7:12, 3] + np.array([-0.1, 0.0, 5e-5, 3e-4, 0.2], np.float32)  # around rt = t
    net.set_input_normalisation(D.to_dev(x_np), rt_mode=rt_mode, t_index=3)
    xn = net.kern.xnorm.cpu().numpy()
    p = net.params.cpu().numpy()
    y = net(D.to_dev(x_np)).cpu().numpy()
    yb, _ = mlp_ref.forward(p, x_np, bf16=True, xnorm=xn, carry="sigmoid")
    assert np.max(np.abs(y - yb)) < 5e-3
    yt = net.torch_reference().cuda()(D.to_dev(x_np))[:, 0].detach().cpu().numpy()
    y64, _ = mlp_ref.forward(p, x_np, xnorm=xn)
    assert _rel(yb, y64) < 4e-2 and abs(_rel(y, yt) - _rel(yb, y64)) < 5e-3
    g = net.input_gradient(D.to_dev(x_np)).cpu().numpy()
    gb, _ = mlp_ref.backward(p, x_np, np.ones(2000), bf16=True, xnorm=xn)
    assert _rel(g[:, :7], gb) < 1e-2
    # the chain rule through the rt transform against central differences of the float64 net
    rows = np.arange(12, 28)
    h = 1e-3
    for j in (3, 5):
        e = np.zeros_like(x_np[rows], dtype=np.float64)
        e[:, j] = h
        fp, _ = mlp_ref.forward(p, (x_np[rows] + e).astype(np.float32), xnorm=xn)
        fm, _ = mlp_ref.forward(p, (x_np[rows] - e).astype(np.float32), xnorm=xn)
        assert _rel(g[rows, j], (fp - fm) / (2 * h)) < 3e-2, j


def test_theta_gradient_matches_finite_differences(torch_gpu):
    from ssmhip import _device as D
    from ssmhip.surrogate import MLPSurrogate

    net = MLPSurrogate(seed=4)
    x = _inputs(16, seed=4)
    g = net.input_gradient(D.to_dev(x)).cpu().numpy()
    p = net.params.cpu().numpy()
    h = 1e-3
    fd = np.zeros((16, 7))
    for j in range(7):
        e = np.zeros_like(x, dtype=np.float64)
        e[:, j] = h
        fp, _ = mlp_ref.forward(p, (x + e).astype(np.float32))
        fm, _ = mlp_ref.forward(p, (x - e).astype(np.float32))
        fd[:, j] = (fp - fm) / (2 * h)
    assert _rel(g[:, :5], fd[:, :5]) < 3e-2


def test_training_reduces_huber_loss_and_checkpoint_roundtrip(torch_gpu):
    from ssmhip import _device as D
    from ssmhip.surrogate import MLPSurrogate

    torch = torch_gpu
    x = _inputs(65536, seed=5)
    target = (np.sin(x[:, 0]) + 0.5 * x[:, 1] - 0.3 * np.log1p(x[:, 5]) * x[:, 6]).astype(np.float32)
    rows = D.to_dev(np.concatenate([x, target[:, None]], axis=1))
    net = MLPSurrogate(seed=5)
    hist = net.fit(rows, epochs=40, batch=4096, lr=3e-3, lr_final=3e-4)
    assert hist[-1] < 0.1 * hist[0]
    pred = net(rows[:, :7].contiguous()).cpu().numpy()
    assert np.mean(np.abs(pred - target)) < 0.1
    sd = net.state_dict()
    other = MLPSurrogate(seed=99)
    other.load_state_dict(sd)
    assert torch.equal(other(rows[:100, :7].contiguous()), net(rows[:100, :7].contiguous()))


def test_train_steps_match_single_steps(torch_gpu):
    """The host-enqueued run of Adam steps (weight gradient summed over batch slices) against
    the same steps taken one call at a time (one-slice weight gradient)."""
    from ssmhip import _device as D
    from ssmhip import _lib as L
    from ssmhip.surrogate import MLPSurrogate

    torch = torch_gpu
    batch, steps = 4096, 6
    x = _inputs(batch * steps, seed=8)
    y = (np.cos(x[:, 0]) - 0.2 * x[:, 1] * x[:, 6]).astype(np.float32)
    dx, dy = D.to_dev(x), D.to_dev(y)
    lrs = np.geomspace(3e-3, 1e-3, steps).astype(np.float32)
    a = MLPSurrogate(seed=8)
    for i in range(steps):
        a.train_step(dx[i * batch:(i + 1) * batch], dy[i * batch:(i + 1) * batch], lr=float(lrs[i]))
    b = MLPSurrogate(seed=8)
    lib = L.load()
    ws_bytes = int(lib.ssm_mlp_train_workspace_bytes(batch))
    assert int(lib.ssm_mlp_train_workspace_bytes(100)) == -1
    ws = D.empty((ws_bytes // 4 + 4,), torch.float32)
    loss = D.empty((batch * steps,), torch.float32)
    L.check(lib.ssm_mlp_train_steps(D.ptr(dx), D.ptr(dy), batch, steps, 0, lrs.ctypes.data, 0.9, 0.999, 1e-8, 0.0,
                                    1.0, D.ptr(b.params), D.ptr(b.m), D.ptr(b.v), D.ptr(b.kern.frags),
                                    D.ptr(b.kern.bias), None, D.ptr(ws), ws_bytes, D.ptr(loss), D.stream_ptr()))
    torch.cuda.synchronize()
    pa, pb = a.params.cpu().numpy(), b.params.cpu().numpy()
    p0 = MLPSurrogate(seed=8).params.cpu().numpy()
    assert _rel(pb - p0, pa - p0) < 1e-3  # the updates agree to float32 summation order
    ya = a(dx[:1000]).cpu().numpy()
    yb = b(dx[:1000]).cpu().numpy()
    assert np.max(np.abs(ya - yb)) < 1e-3
    assert np.isfinite(loss.cpu().numpy()).all()
    # the fused Adam + repack leaves exactly the packing of the updated parameters
    fr = torch.zeros_like(b.kern.frags)
    bi = torch.zeros_like(b.kern.bias)
    L.check(lib.ssm_mlp_pack(D.ptr(b.params), D.ptr(fr), D.ptr(bi), D.stream_ptr()))
    torch.cuda.synchronize()
    assert torch.equal(fr, b.kern.frags) and torch.equal(bi, b.kern.bias)
    with pytest.raises(ValueError):
        L.check(lib.ssm_mlp_train_steps(D.ptr(dx), D.ptr(dy), 100, 1, 0, lrs.ctypes.data, 0.9, 0.999, 1e-8, 0.0,
                                        1.0, D.ptr(b.params), D.ptr(b.m), D.ptr(b.v), D.ptr(b.kern.frags),
                                        D.ptr(b.kern.bias), None, D.ptr(ws), ws_bytes, None, D.stream_ptr()))


@pytest.mark.parametrize("mode", [0, 1])
def test_team_kernel_matches_single_wave_kernel_bit_for_bit(torch_gpu, mode):
    """Without dx the library runs four waves per 16-row tile; with dx one wave per tile.
    Both accumulate every element in the same order, so all shared outputs agree exactly."""
    from ssmhip import _device as D
    from ssmhip import _lib as L
    from ssmhip.surrogate import MLPSurrogate

    torch = torch_gpu
    net = MLPSurrogate(seed=3)
    n, ld = 1000, 1024
    x_np = _inputs(n, seed=9)
    dxin = D.to_dev(x_np)
    net.set_input_normalisation(dxin)
    tgt = D.to_dev(np.random.default_rng(4).normal(size=n).astype(np.float32))
    sz = net.kern.sz
    lib = L.load()
    outs = []
    for want_dx in (False, True):
        acts = D.zeros((sz["act_rows"], ld), torch.bfloat16)
        deltas = D.zeros((sz["delta_rows"], ld), torch.bfloat16)
        y = D.zeros((n,), torch.float32)
        loss = D.zeros((n,), torch.float32)
        dx = D.zeros((n, 8), torch.float32) if want_dx else None
        L.check(lib.ssm_mlp_backward(D.ptr(dxin), D.ptr(tgt), n, mode, 1.0, 1.0 / n, D.ptr(net.kern.frags),
                                     D.ptr(net.kern.bias), D.ptr(net.kern.xnorm), D.ptr(acts), D.ptr(deltas), ld,
                                     D.ptr(dx), D.ptr(y), D.ptr(loss), D.stream_ptr()))
        torch.cuda.synchronize()
        outs.append([t.float().cpu().numpy() for t in (y, loss, acts, deltas)])
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    assert np.abs(outs[0][3]).sum() > 0
