"""Set records laid out on the device from per-slot columns (ssm_em_records) equal the host
layout of the same columns bit for bit, for every single-particle diffusion record kind, at set
counts on either side of the one from which the simulators lay records out on the device; the
uniform offsets written on the device equal k * n_samples."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


def _cases(n, rng):
    from ssmhip import _lib as L

    v = rng.uniform(-3, 3, n).astype(F32)
    a = rng.uniform(0.3, 3, n).astype(F32)
    z = rng.uniform(0.1, 0.9, n).astype(F32)
    t = rng.uniform(0, 2, n).astype(F32)
    dl = np.where(rng.uniform(size=n) < 0.5, 999.0, rng.uniform(0.5, 5, n)).astype(F32)
    s = rng.uniform(0.5, 1.5, n).astype(F32)
    base = (v, a, z, t, dl, s, 20.0, 0.001)
    yield "ddm angle", base + (L.BOUNDARY_ANGLE, {"a": a, "theta": rng.uniform(-0.1, 1.3, n).astype(F32)}), {}
    yield "ddm weibull", base + (L.BOUNDARY_WEIBULL, {"a": a, "alpha": rng.uniform(0.3, 5, n).astype(F32),
                                                       "beta": rng.uniform(0.3, 7, n).astype(F32)}), {}
    yield "full ddm", base + (L.BOUNDARY_CONSTANT, {"a": a}), {"sz": rng.uniform(0, 0.3, n), "sv": rng.uniform(0, 1, n),
                                                              "st": rng.uniform(0, 0.2, n)}
    yield "ornstein", base + (L.BOUNDARY_CONSTANT, {"a": a}), {"g": rng.uniform(-1, 1, n)}
    yield "levy", base + (L.BOUNDARY_CONSTANT, {"a": a}), {"alpha": rng.uniform(1.0, 2.0, n)}


@pytest.mark.parametrize("n", [1, 257, 4095, 4096, 100_003])
def test_device_records_equal_host_records(torch_gpu, n):
    from ssmhip import _lib as L
    from ssmhip import em

    rng = np.random.default_rng(n)
    for name, args, kw in _cases(n, rng):
        print(f"[case] records on the device: {name}, {n} sets", flush=True)
        for stride in (L.SET_STRIDE_SINGLE, L.SET_STRIDE):
            host = em.ddm_records(*args, **kw, stride=stride)
            # the device layout itself, and the helper that picks a side by the set count
            n_cols, cols = em.ddm_columns(*args, **kw)
            for dev in (em.records_on_device(n_cols, cols, stride).cpu().numpy(),
                        em.ddm_records_device(*args, **kw, stride=stride).cpu().numpy()):
                assert dev.shape == host.shape
                assert np.array_equal(dev.view(np.uint32), host.view(np.uint32)), name


def test_device_records_for_zero_sets_and_uniform_offsets(torch_gpu):
    from ssmhip import _lib as L
    from ssmhip import em

    empty = np.zeros(0, F32)
    rec = em.ddm_records_device(empty, empty, empty, empty, empty, empty, 20.0, 0.001, L.BOUNDARY_CONSTANT,
                                {"a": empty})
    assert tuple(rec.shape) == (0, L.SET_STRIDE_SINGLE)
    for n_sets, n in ((0, 5), (1, 0), (3, 7), (1_000_000, 1), (4097, 100_000)):
        got = em.uniform_offsets_device(n_sets, n).cpu().numpy()
        assert np.array_equal(got, em.uniform_offsets(n_sets, n))


def test_simulator_outputs_unchanged_by_device_records(torch_gpu):
    """The Simulator path builds its records on the device: a replayed DDM run through it
    equals the oracle transcription trajectory by trajectory."""
    from oracle import em_numpy as E
    from ssmhip import cssm

    rng = np.random.default_rng(5)
    K, N, stride = 9, 40, 1031
    v = rng.uniform(-2, 2, K).astype(F32)
    a = rng.uniform(0.5, 2.0, K).astype(F32)
    z = rng.uniform(0.2, 0.8, K).astype(F32)
    t = rng.uniform(0.0, 1.0, K).astype(F32)
    th = rng.uniform(-0.1, 1.0, K).astype(F32)
    normals = rng.standard_normal((N * K, stride)).astype(F32)
    unif = rng.uniform(size=N * K).astype(F32)
    from ssmhip.boundary_functions import angle

    out = cssm.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, n_trials=K,
                             boundary_fun=angle, boundary_params={"theta": th}, smooth_unif=True,
                             return_option="minimal", replay_normals=normals, replay_uniforms=unif)
    # cssm returns [n_samples, n_trials, 1]; the replay rows are set-major (trial k, sample n)
    r0, c0, _ = E.ddm_flexbound(v, a, z, t, np.full(K, 999, F32), np.ones(K, F32), n_samples=N, delta_t=0.001,
                                boundary="angle", bparams={"a": a, "theta": th}, smooth_unif=True,
                                normals=normals, uniforms=unif)
    got_r = out["rts"][:, :, 0].T.reshape(-1)
    got_c = out["choices"][:, :, 0].T.reshape(-1)
    assert np.array_equal(got_c, c0)
    assert np.array_equal(got_r, r0)


def test_host_copies_equal_plain_copies_on_each_side_of_the_pinned_range(torch_gpu):
    """Simulator outputs reach the host through pinned buffers from 1 MiB to 64 MiB and through
    ordinary copies outside that range; every path gives the device values."""
    import torch

    from ssmhip import _device as D

    for n in (1000, 300_000, 17_000_000):  # 4 KB, 1.2 MB, 68 MB of float32
        print(f"[case] host copy of {n} floats", flush=True)
        x = torch.arange(n, dtype=torch.float32, device="cuda") * 0.5
        y = (torch.arange(n, dtype=torch.int32, device="cuda") % 7) - 3
        a, b = D.to_host(x, y)
        assert a.dtype == np.float32 and b.dtype == np.int32
        assert np.array_equal(a, x.cpu().numpy()) and np.array_equal(b, y.cpu().numpy())
        del x, y
