"""NumPy reference of the surrogate MLP (test oracle).

7 -> 100 -> 100 -> 120 -> 1 with tanh hidden layers, nn.Linear parameter
layout, flat parameter vector [W1, b1, W2, b2, W3, b3, W4, b4].
`bf16=True` rounds every GEMM operand (inputs, weights, hidden activations,
backward deltas) to bfloat16 (round to nearest even) and accumulates in
float64, i.e. the arithmetic the MFMA kernels perform up to summation order;
the tanh derivative 1 - h^2 then uses the bf16 activations, as the kernels do,
and hidden-layer forward weights are rounded as bf16(2 log2(e) W) (the scale
the kernels fold into their packed weights). `carry="sigmoid"` models the
inference kernel, which carries each hidden unit as u = (1 - tanh z) / 2 and
folds 1 - 2u into the next layer: with bf16 rounding the next layer then sees
1 - 2 bf16(u) in place of bf16(tanh z); in exact arithmetic the two agree.
Pinned by: PyTorch autograd of the same nn.Sequential (float64) and central
finite differences (tests/test_oracle_mlp.py).
"""

from __future__ import annotations

import numpy as np

LAYERS = [(100, 7), (100, 100), (120, 100), (1, 120)]
TANH_SCALE = 2.8853900817779268  # 2 log2(e)


def unflatten(p: np.ndarray):
    out, off = [], 0
    for o, n in LAYERS:
        W = p[off:off + o * n].reshape(o, n)
        off += o * n
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    u = (u + np.uint64(0x7FFF) + lsb) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def rt_feature(x: np.ndarray, xnorm) -> np.ndarray | None:
    """u with input 5 = log(max(u, 1e-4)) (float32): u = rt (mode 1) or rt - x[t index]
    (mode 2); None for mode 0."""
    xn = np.asarray(xnorm, np.float32)
    mode = float(xn[14]) if xn.shape[0] > 14 else 0.0
    if mode == 0.0:
        return None
    u = np.array(x, np.float32)[:, 5]
    if mode == 2.0:
        u = (u - np.array(x, np.float32)[:, int(xn[15])]).astype(np.float32)
    return u


def standardise(x: np.ndarray, xnorm) -> np.ndarray:
    """(x' - shift) * scale in float32; x' = x with input 5 replaced by log(max(u, 1e-4)),
    u from rt_feature, when xnorm[14] (the rt mode) is 1 or 2."""
    x = np.array(x, np.float32)
    if xnorm is None:
        return x
    xn = np.asarray(xnorm, np.float32)
    u = rt_feature(x, xnorm)
    if u is not None:
        x[:, 5] = np.log(np.maximum(u, np.float32(1e-4)))
    return ((x - xn[:7]).astype(np.float32) * xn[7:14]).astype(np.float32)


def forward(p: np.ndarray, x: np.ndarray, bf16: bool = False, xnorm=None, carry: str = "tanh"):
    """Returns y [n] and the list of hidden activations."""
    if carry not in ("tanh", "sigmoid"):
        raise ValueError(carry)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(standardise(x, xnorm))
    hs = [h]
    layers = unflatten(np.asarray(p, np.float32))
    for i, (W, b) in enumerate(layers):
        if bf16 and i < 3:
            # hidden-layer weights are stored as bf16(2 log2(e) W): the forward kernels fold
            # the tanh argument scale into the packed weights
            Wr = bf16_round((np.float32(TANH_SCALE) * W).astype(np.float32)) / TANH_SCALE
        else:
            Wr = r(W)
        z = h @ Wr.T + np.asarray(b, np.float64)
        if i < 3:
            a = np.tanh(z)
            hs.append(a)
            if bf16 and carry == "sigmoid":
                h = 1.0 - 2.0 * r((0.5 * (1.0 - a)).astype(np.float32))
            else:
                h = r(a.astype(np.float32)) if bf16 else a
        else:
            return z[:, 0], hs
    raise AssertionError


def backward(p: np.ndarray, x: np.ndarray, gy: np.ndarray, bf16: bool = False, xnorm=None):
    """dL/dx [n, 7] (raw inputs) and dL/dp (flat) for upstream gradient gy = dL/dy."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(p, np.float32))
    _, hs = forward(p, x, bf16, xnorm)
    # operands the kernels feed to the MFMA
    ins = [r(standardise(x, xnorm))] + [r(h.astype(np.float32)) if bf16 else h for h in hs[1:]]
    dz = np.asarray(gy, np.float64)[:, None]
    grads = [None] * 4
    for i in range(3, -1, -1):
        W, _ = layers[i]
        dzr = r(dz.astype(np.float32)) if bf16 else dz
        grads[i] = (dzr.T @ ins[i], dzr.sum(0))
        dh = dzr @ r(W)
        if i > 0:
            h = ins[i]  # the kernels take the derivative from the bf16 activations they keep
            dz = dh * (1.0 - h ** 2)
        else:
            dx = dh
    flat = np.concatenate([np.concatenate([g[0].reshape(-1), g[1]]) for g in grads])
    if xnorm is not None:
        xn = np.asarray(xnorm, np.float64)
        dx = dx * xn[7:14]
        u = rt_feature(x, xnorm)
        if u is not None:
            u = u.astype(np.float64)
            dx[:, 5] = np.where(u > 1e-4, dx[:, 5] / np.maximum(u, 1e-4), 0.0)
            if xn[14] == 2:
                dx[:, int(xn[15])] -= dx[:, 5]
    return dx, flat


def huber(e: np.ndarray, delta: float) -> np.ndarray:
    a = np.abs(e)
    return np.where(a <= delta, 0.5 * e * e, delta * (a - 0.5 * delta))
