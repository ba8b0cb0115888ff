"""Learned likelihood surrogate: an MLP 7 -> 100 -> 100 -> 120 -> 1 (tanh) over
LAN rows (5 angle-DDM theta, rt, choice) -> KDE log-likelihood, with
hand-written BF16-MFMA forward / backward kernels (FP32 accumulation, FP32
master weights) behind a ``torch.autograd.Function``.

The reference generates the training rows (ResampleMixtureStrategy,
ssms/dataset_generators/strategies/resample_mixture_strategy.py:64-214) but
has no network; the KDE path stays the fallback outside the training bounds
and the check of the surrogate's error (``SurrogateLikelihood``).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _device as D
from . import _lib as L

IN, H1, H2, H3, OUT = 7, 100, 100, 120, 1
RT_MODES = {"raw": 0, "log": 1, "log_minus_t": 2}  # the rt input transform (xnorm[14])
LABEL_LOWER_BOUND = math.log(1e-7)  # training-target floor of fit()
FLOOR = -66.774  # the KDE labels' log-likelihood floor (kde_class.py:188-195)
LAYERS = [(H1, IN), (H2, H1), (H3, H2), (OUT, H3)]


def sizes() -> dict:
    vals = [C.c_int32() for _ in range(5)]
    L.check(L.load().ssm_mlp_sizes(*[C.byref(v) for v in vals]), "ssm_mlp_sizes")
    keys = ("n_params", "n_frag", "n_bias", "act_rows", "delta_rows")
    return {k: int(v.value) for k, v in zip(keys, vals)}


def param_slices() -> list[tuple[str, slice, tuple]]:
    """(name, slice into the flat buffer, shape) in nn.Linear layout."""
    out, off = [], 0
    for i, (o, n) in enumerate(LAYERS, 1):
        out.append((f"W{i}", slice(off, off + o * n), (o, n)))
        off += o * n
        out.append((f"b{i}", slice(off, off + o), (o,)))
        off += o
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """nn.Linear's default init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, n in LAYERS:
        bound = 1.0 / math.sqrt(n)
        parts.append(rng.uniform(-bound, bound, o * n))
        parts.append(rng.uniform(-bound, bound, o))
    return np.concatenate(parts).astype(np.float32)


class _Kernels:
    """Packed weights on the device plus thin launch helpers."""

    def __init__(self):
        t = D.require_gpu()
        self.sz = sizes()
        self.frags = D.empty((self.sz["n_frag"],), t.bfloat16)
        self.bias = D.empty((self.sz["n_bias"],), t.float32)
        self.xnorm = None  # device float32[16]: shift[7], scale[7], rt mode, t index; or None

    def pack(self, params) -> None:
        L.check(L.load().ssm_mlp_pack(D.ptr(params), D.ptr(self.frags), D.ptr(self.bias), D.stream_ptr()),
                "ssm_mlp_pack")

    def forward(self, x):
        t = D.torch()
        n = x.shape[0]
        y = D.empty((max(n, 1),), t.float32)
        L.check(L.load().ssm_mlp_forward(D.ptr(x), n, D.ptr(self.frags), D.ptr(self.bias), D.ptr(self.xnorm),
                                         D.ptr(y), D.stream_ptr()), "ssm_mlp_forward")
        return y[:n]

    def backward(self, x, target_or_grad, mode: int, *, delta=1.0, scale=1.0, want_dx=True, want_w=True):
        t = D.torch()
        n = x.shape[0]
        ld = max(32, (n + 31) // 32 * 32)
        acts = D.zeros((self.sz["act_rows"], ld), t.bfloat16) if want_w else None
        deltas = D.zeros((self.sz["delta_rows"], ld), t.bfloat16) if want_w else None
        dx = D.empty((n, 8), t.float32) if want_dx else None
        y = D.empty((n,), t.float32)
        loss = D.empty((n,), t.float32) if mode == 1 else None
        rc = L.load().ssm_mlp_backward(D.ptr(x), D.ptr(target_or_grad), n, mode, float(delta), float(scale),
                                       D.ptr(self.frags), D.ptr(self.bias), D.ptr(self.xnorm), D.ptr(acts),
                                       D.ptr(deltas), ld,
                                       D.ptr(dx), D.ptr(y), D.ptr(loss), D.stream_ptr())
        L.check(rc, "ssm_mlp_backward")
        grad = None
        if want_w:
            grad = D.zeros((self.sz["n_params"],), t.float32)
            L.check(L.load().ssm_mlp_wgrad(D.ptr(acts), D.ptr(deltas), n, ld, D.ptr(grad), D.stream_ptr()),
                    "ssm_mlp_wgrad")
        return y, dx, grad, loss


def _autograd_function():
    torch = D.torch()

    class MLPFunction(torch.autograd.Function):
        """y = net(x; params) on the MFMA kernels; backward w.r.t. x and params by the hand-written kernels."""

        @staticmethod
        def forward(ctx, x, params, kern):
            x = x.contiguous()
            kern.pack(params)
            y = kern.forward(x)
            ctx.save_for_backward(x, params)
            ctx.kern = kern
            return y

        @staticmethod
        def backward(ctx, grad_y):
            x, params = ctx.saved_tensors
            kern = ctx.kern
            kern.pack(params)
            _, dx, dparams, _ = kern.backward(x, grad_y.contiguous().float(), 0,
                                              want_dx=ctx.needs_input_grad[0], want_w=ctx.needs_input_grad[1])
            return (dx[:, :IN] if dx is not None else None), dparams, None

    return MLPFunction


_FN = None


def mlp_apply(x, params, kern=None):
    """Differentiable y = net(x; params) (x: [n, 7] float32 cuda, params: flat float32 cuda)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(x, params, kern or _Kernels())


class MLPSurrogate:
    """FP32 master weights + Adam state on the device; BF16-MFMA compute."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None, rt_mode: str = "log_minus_t",
                 t_index: int = 3):
        t = D.require_gpu()
        if rt_mode not in RT_MODES:
            raise ValueError(f"rt_mode must be one of {sorted(RT_MODES)}")
        self.rt_mode, self.t_index = rt_mode, int(t_index)
        self.kern = _Kernels()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape[0] != self.kern.sz["n_params"]:
            raise ValueError(f"expected {self.kern.sz['n_params']} parameters, got {p.shape[0]}")
        self.params = D.to_dev(p)
        self.m = D.zeros(tuple(self.params.shape), t.float32)
        self.v = D.zeros(tuple(self.params.shape), t.float32)
        self.step = 0
        self.kern.pack(self.params)
        self.bounds = None  # (lo[7], hi[7]) of the training inputs
        self.val_mae = None  # mean |error| vs held-out KDE labels above the floor
        self.val_mae_all = None

    # ---- inference -----------------------------------------------------------
    def __call__(self, x):
        return self.kern.forward(x.contiguous())

    def input_gradient(self, x):
        """d y / d x for every row (e.g. d loglik / d theta = [:, :5])."""
        t = D.torch()
        ones = t.ones((x.shape[0],), dtype=t.float32, device=x.device)
        _, dx, _, _ = self.kern.backward(x.contiguous(), ones, 0, want_dx=True, want_w=False)
        return dx[:, :IN]

    def differentiable(self, x):
        return mlp_apply(x, self.params, self.kern)

    # ---- training ------------------------------------------------------------
    def train_step(self, x, target, *, lr=1e-3, delta=1.0, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        """One Adam step on the Huber loss over the batch (fused kernels); returns the batch loss tensor."""
        n = x.shape[0]
        _, _, grad, loss = self.kern.backward(x, target, 1, delta=delta, scale=1.0 / n, want_dx=False, want_w=True)
        self.step += 1
        L.check(L.load().ssm_mlp_adam(D.ptr(self.params), D.ptr(grad), D.ptr(self.m), D.ptr(self.v), self.step,
                                      float(lr), float(betas[0]), float(betas[1]), float(eps), float(weight_decay),
                                      D.stream_ptr()), "ssm_mlp_adam")
        self.kern.pack(self.params)
        return loss

    def fit(self, rows, *, epochs=10, batch=8192, lr=1e-3, lr_final=1e-4, delta=1.0, seed=0, val_rows=None,
            label_lower_bound=LABEL_LOWER_BOUND, log=None):
        """Train on LAN rows [N, 8] (device tensor: 7 inputs + loglik label). The training
        targets are clipped from below at `label_lower_bound` (log(1e-7) by default, None: no
        clipping): the floor rows (-66.77) then no longer pull the fit away from the
        likelihood's bulk, where its accuracy matters."""
        t = D.torch()
        rows = rows.contiguous()
        n = rows.shape[0]
        if self.kern.xnorm is None:
            self.set_input_normalisation(rows[:, :IN], self.rt_mode, self.t_index)
        if batch % 32:
            raise ValueError("batch must be a multiple of 32")
        steps_per_epoch = max(1, n // batch)
        if n < batch:
            raise ValueError(f"need at least one batch of {batch} rows, got {n}")
        total = epochs * steps_per_epoch
        gen = t.Generator(device=rows.device)
        gen.manual_seed(seed)
        lib = L.load()
        ws_bytes = int(lib.ssm_mlp_train_workspace_bytes(batch))
        ws = D.empty((ws_bytes // 4 + 4,), t.float32)
        loss = D.empty((steps_per_epoch * batch,), t.float32)
        history = []
        for ep in range(epochs):
            perm = t.randperm(n, device=rows.device, generator=gen)[: steps_per_epoch * batch]
            shuffled = rows.index_select(0, perm)
            x_all = shuffled[:, :IN].contiguous()
            y_all = shuffled[:, IN].contiguous()
            if label_lower_bound is not None:
                y_all.clamp_(min=float(label_lower_bound))
            k = np.arange(ep * steps_per_epoch, (ep + 1) * steps_per_epoch, dtype=np.float64)
            lrs = (lr * (lr_final / lr) ** (k / max(1, total - 1))).astype(np.float32)  # exponential decay
            L.check(lib.ssm_mlp_train_steps(D.ptr(x_all), D.ptr(y_all), batch, steps_per_epoch, self.step,
                                            lrs.ctypes.data, 0.9, 0.999, 1e-8, 0.0, float(delta), D.ptr(self.params),
                                            D.ptr(self.m), D.ptr(self.v), D.ptr(self.kern.frags),
                                            D.ptr(self.kern.bias), D.ptr(self.kern.xnorm), D.ptr(ws), ws_bytes,
                                            D.ptr(loss), D.stream_ptr()), "ssm_mlp_train_steps")
            self.step += steps_per_epoch
            history.append(float(loss.sum().item()) / (steps_per_epoch * batch))  # mean per-row Huber loss
            if log:
                log(f"epoch {ep}: huber {history[-1]:.5f}")
        xs = rows[:, :IN]
        self.bounds = (xs.min(0).values.cpu().numpy(), xs.max(0).values.cpu().numpy())
        if val_rows is not None:
            pred = self(val_rows[:, :IN].contiguous())
            err = (pred - val_rows[:, IN]).abs()
            above = val_rows[:, IN] > FLOOR + 1.0  # rows the KDE labels above its floor
            self.val_mae_all = float(err.mean().item())
            self.val_mae = float(err[above].mean().item()) if bool(above.any()) else self.val_mae_all
        return history

    # ---- inference over theta ---------------------------------------------------
    def fit_theta(self, rt, choice, *, lower=None, upper=None, n_starts: int = 16, steps: int = 300,
                  lr: float = 0.05, lr_final: float = 0.002, seed: int = 0):
        """Maximum-likelihood parameters of one data set under the surrogate.

        Gradient ascent on sum_i loglik(theta, rt_i, choice_i) over the 5 parameter inputs:
        d loglik / d theta comes from the hand-written backward kernel through the autograd
        function (`differentiable`), summed over the data rows by autograd. Adam runs
        `n_starts` starts at once, drawn uniformly in [lower, upper] (default: the training
        box of the parameter inputs), with an exponential learning-rate decay lr -> lr_final.
        After every step the parameters are projected back into the box.

        A trial with rt <= t (the non-decision time input, column `t_index`) is impossible under
        the model: it contributes the KDE labels' floor (FLOOR, -66.774) instead of the
        network's value, as the KDE labels such rows. Without that rule the fit can drift into a
        corner of the box where the network, trained on floor-clipped targets, extrapolates.

        Returns (theta_hat [5], loglik_hat, thetas [n_starts, 5], logliks [n_starts]) as NumPy
        arrays; theta_hat is the start with the largest final log-likelihood."""
        t = D.torch()
        if lower is None or upper is None:
            if self.bounds is None:
                raise ValueError("fit_theta needs lower/upper or a trained surrogate (its training bounds)")
        lo = np.asarray(self.bounds[0][:5] if lower is None else lower, np.float32)
        hi = np.asarray(self.bounds[1][:5] if upper is None else upper, np.float32)
        rt = t.as_tensor(np.asarray(rt, np.float32).reshape(-1), device="cuda")
        ch = t.as_tensor(np.asarray(choice, np.float32).reshape(-1), device="cuda")
        if rt.numel() == 0 or rt.numel() != ch.numel():
            raise ValueError("rt and choice must be non-empty and of the same length")
        K, N = int(n_starts), int(rt.numel())
        rng = np.random.default_rng(seed)
        lo_t, hi_t = t.as_tensor(lo, device="cuda"), t.as_tensor(hi, device="cuda")
        theta = t.tensor(rng.uniform(lo, hi, (K, 5)).astype(np.float32), device="cuda", requires_grad=True)
        opt = t.optim.Adam([theta], lr=lr)
        gamma = (lr_final / lr) ** (1.0 / max(1, steps))
        sched = t.optim.lr_scheduler.ExponentialLR(opt, gamma=gamma)
        data = t.stack([rt, ch], 1)[None].expand(K, N, 2)

        def rows(th):
            return t.cat([th[:, None, :].expand(K, N, 5), data], 2).reshape(K * N, IN)

        def total(y, th):
            possible = rt[None, :] > th[:, self.t_index:self.t_index + 1]
            return t.where(possible, y.view(K, N), t.full_like(y.view(K, N), FLOOR)).sum(1)

        for _ in range(int(steps)):
            ll = total(self.differentiable(rows(theta)), theta)
            opt.zero_grad(set_to_none=True)
            (-ll.sum()).backward()
            opt.step()
            sched.step()
            with t.no_grad():
                theta.copy_(t.minimum(t.maximum(theta, lo_t), hi_t))
        with t.no_grad():
            ll = total(self(rows(theta)), theta)
        best = int(t.argmax(ll).item())
        thetas, lls = theta.detach().cpu().numpy(), ll.cpu().numpy()
        return thetas[best], float(lls[best]), thetas, lls

    def set_input_normalisation(self, x, rt_mode: str = "log_minus_t", t_index: int = 3) -> None:
        """Standardise the inputs: the network sees (x' - mean) / std (std floored at 1e-6).
        x' is x with rt replaced by log(max(u, 1e-4)), u = rt ("log": the KDE labels live on
        log rt) or u = rt - x[t_index] ("log_minus_t": the time since the non-decision time,
        where the density has its onset); "raw" leaves rt as it is."""
        t = D.torch()
        if rt_mode not in RT_MODES:
            raise ValueError(f"rt_mode must be one of {sorted(RT_MODES)}")
        if not 0 <= int(t_index) < 5:
            raise ValueError("t_index must name one of the 5 parameter inputs")
        xx = x.float().clone()
        if rt_mode == "log_minus_t":
            xx[:, 5] = xx[:, 5] - xx[:, int(t_index)]
        if rt_mode != "raw":
            xx[:, 5] = t.log(t.clamp_min(xx[:, 5], 1e-4))
        mu = xx.mean(0)
        sd = xx.std(0).clamp_min(1e-6)
        tail = t.tensor([float(RT_MODES[rt_mode]), float(t_index)], device=x.device)
        self.kern.xnorm = t.cat([mu, 1.0 / sd, tail]).float().contiguous()
        self.rt_mode, self.t_index = rt_mode, int(t_index)

    # ---- checkpoint ----------------------------------------------------------
    def state_dict(self) -> dict:
        d = {"params": self.params.detach().cpu(), "adam_m": self.m.cpu(), "adam_v": self.v.cpu(), "step": self.step}
        if self.kern.xnorm is not None:
            d["xnorm"] = self.kern.xnorm.cpu()
        if self.bounds is not None:
            d["bounds_lo"], d["bounds_hi"] = self.bounds
        return d

    def load_state_dict(self, d: dict) -> None:
        self.params.copy_(d["params"].to(self.params.device))
        self.m.copy_(d["adam_m"].to(self.m.device))
        self.v.copy_(d["adam_v"].to(self.v.device))
        self.step = int(d["step"])
        if "xnorm" in d:
            self.kern.xnorm = d["xnorm"].to(self.params.device).float().contiguous()
        if "bounds_lo" in d:
            self.bounds = (np.asarray(d["bounds_lo"]), np.asarray(d["bounds_hi"]))
        self.kern.pack(self.params)

    def torch_reference(self):
        """The same network as a torch.nn.Sequential (float32) with the current master weights."""
        torch = D.torch()
        layers = []
        if self.kern.xnorm is not None:
            xn = self.kern.xnorm.detach().cpu()

            class Standardise(torch.nn.Module):
                def __init__(self):
                    super().__init__()
                    self.register_buffer("shift", xn[:IN].clone())
                    self.register_buffer("scale", xn[IN:2 * IN].clone())
                    self.mode = int(xn[2 * IN]) if xn.shape[0] > 2 * IN else 0
                    self.ti = int(xn[2 * IN + 1]) if xn.shape[0] > 2 * IN + 1 else 0

                def forward(self, x):
                    if self.mode:
                        u = x[:, 5:6] - x[:, self.ti:self.ti + 1] if self.mode == 2 else x[:, 5:6]
                        x = torch.cat([x[:, :5], torch.log(torch.clamp_min(u, 1e-4)), x[:, 6:]], dim=1)
                    return (x - self.shift) * self.scale

            layers.append(Standardise())
        net = torch.nn.Sequential(*layers, torch.nn.Linear(IN, H1), torch.nn.Tanh(), torch.nn.Linear(H1, H2),
                                  torch.nn.Tanh(), torch.nn.Linear(H2, H3), torch.nn.Tanh(), torch.nn.Linear(H3, OUT))
        p = self.params.detach().cpu()
        lin = [m for m in net if isinstance(m, torch.nn.Linear)]
        sl = param_slices()
        with torch.no_grad():
            for i, layer in enumerate(lin):
                _, ws, wshape = sl[2 * i]
                _, bs, _ = sl[2 * i + 1]
                layer.weight.copy_(p[ws].reshape(wshape))
                layer.bias.copy_(p[bs])
        return net


class SurrogateLikelihood:
    """Surrogate log-likelihood with the KDE path as fallback and as the check of its error.

    Rows outside the surrogate's training input box (or every row, when the
    validation error exceeds `max_val_mae`) are labelled by simulating their
    parameter set and evaluating the GPU KDE at the row.
    """

    def __init__(self, surrogate: MLPSurrogate, model: str = "angle", max_val_mae: float = 1.0,
                 kde_samples: int = 100_000, seed: int = 0):
        self.net, self.model, self.max_val_mae = surrogate, model, max_val_mae
        self.kde_samples, self.seed = kde_samples, seed
        self.fallback_rows = 0

    def in_bounds(self, x_np: np.ndarray) -> np.ndarray:
        if self.net.bounds is None:
            return np.zeros(x_np.shape[0], bool)
        lo, hi = self.net.bounds
        return np.all((x_np >= lo) & (x_np <= hi), axis=1)

    def loglik(self, x):
        """x: device tensor [n, 7] -> loglik [n] (device)."""
        y = self.net(x)
        x_np = x.cpu().numpy()
        use_kde = ~self.in_bounds(x_np)
        if self.net.val_mae is not None and self.net.val_mae > self.max_val_mae:
            use_kde[:] = True
        if use_kde.any():
            y = y.clone()
            idx = np.nonzero(use_kde)[0]
            y[D.to_dev(idx.astype(np.int64))] = D.to_dev(self.kde_loglik(x_np[idx]))
            self.fallback_rows += int(idx.size)
        return y

    def kde_loglik(self, x_np: np.ndarray) -> np.ndarray:
        """GPU KDE labels for rows (theta, rt, choice): one simulated set per distinct theta."""
        from .kde import KDEBank, PointBatch  # noqa: PLC0415
        from .lan import choice_values_for  # noqa: PLC0415
        from .batched import BatchedSimulation  # noqa: PLC0415

        theta = x_np[:, :5].astype(np.float32)
        uniq, inv = np.unique(theta, axis=0, return_inverse=True)
        sim = BatchedSimulation(self.model, uniq, self.kde_samples, seed=self.seed, smooth_unif=True, launch=None)
        sim.run()
        cv = choice_values_for(self.model)
        bank = KDEBank(sim.rts, sim.choices, sim.job.offsets, cv)
        out = np.empty(x_np.shape[0], np.float32)
        lookup = {c: i for i, c in enumerate(cv)}
        for k in range(uniq.shape[0]):
            rows = np.nonzero(inv.reshape(-1) == k)[0]
            ci = np.array([lookup[int(round(c))] for c in x_np[rows, 6]])
            sub = _SubBank(bank, k)
            pts, order = PointBatch.from_host(x_np[rows, 5], ci, 1, len(cv))
            ll = sub.evaluate(pts).cpu().numpy()
            res = np.empty(rows.size, np.float32)
            res[order] = ll
            out[rows] = res
        return out


class _SubBank:
    """View of one set of a KDEBank as a one-set bank (shares device memory)."""

    def __init__(self, bank, k):
        t = D.torch()
        C = bank.C
        self.bank, self.k, self.C = bank, k, C
        g = slice(k * C, (k + 1) * C)
        self.kde_count = bank.kde_count[g].contiguous()
        self.choice_total = bank.choice_total[g].contiguous()
        self.bandwidth = bank.bandwidth[g].contiguous()
        self.kde_start = bank.kde_start[g].contiguous()
        self.offsets = bank.offsets[k:k + 2].contiguous()
        self.max_samples = bank.max_samples

    def evaluate(self, pts, max_t=20.0, lb=-66.774, eps=10e-5):
        t = D.torch()
        out = D.empty(tuple(pts.rt.shape), t.float32)
        nch = max(1, -(-self.max_samples // L.KDE_CHUNK))
        ws = D.empty((int(pts.rt.shape[0]) * nch,), t.float32)
        rc = L.load().ssm_kde_eval(1, self.C, D.ptr(pts.grp_begin), D.ptr(pts.grp_count), D.ptr(pts.rt),
                                   D.ptr(self.bank.samples), D.ptr(self.kde_count), D.ptr(self.choice_total),
                                   D.ptr(self.bandwidth), D.ptr(self.kde_start), D.ptr(self.offsets), float(max_t),
                                   float(lb), float(eps), self.max_samples, D.ptr(ws), D.ptr(out), D.stream_ptr())
        L.check(rc, "ssm_kde_eval")
        return out
