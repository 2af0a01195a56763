"""The learned vertical closure: the FP64 numpy restatement (oracle/mlp.py)
pinned by finite differences, the HIP feature kernel, the HIP training
kernels' gradients, the BF16 MFMA inference against FP32, training, the
per-column MY2.5 fallback, and what the trained network gains over the
closure-free method (constant vertical mixing)."""
import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import mlp


def _rand_params(seed=1, scale=1.0):
    from myfvcomcode_amd.closure import init_params
    return init_params(seed) * scale


def _samples(n=2000, seed=3):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 8), np.float32)
    x[:, 0] = rng.normal(0, 2, n)
    x[:, 1] = rng.normal(0, 1, n)
    x[:, 2] = -rng.uniform(0, 1, n)
    x[:, 3] = rng.uniform(0, 2, n)
    x[:, 4] = rng.normal(0, 0.5, n)
    x[:, 5] = 1.0
    y = np.exp(rng.uniform(np.log(1e-5), np.log(1e-1), (n, 2))).astype(np.float32)
    return x, y


def test_reference_gradient_matches_finite_differences():
    p = _rand_params().astype(np.float64)
    x, y = _samples(64)
    loss, g = mlp.loss_grad(p, x, y)
    rng = np.random.default_rng(0)
    for j in rng.choice(mlp.NPARAM, 40, replace=False):
        h = 1e-6 * max(1.0, abs(p[j]))
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (mlp.loss_grad(pp, x, y)[0] - mlp.loss_grad(pm, x, y)[0]) / (2 * h)
        assert fd == pytest.approx(g[j], rel=1e-4, abs=1e-9), j


def test_reference_forward_is_the_loss_model():
    p = _rand_params().astype(np.float64)
    x, y = _samples(32)
    k = mlp.forward(p, x)
    assert (k > 0).all()
    loss, _ = mlp.loss_grad(p, x, k)       # predictions as targets: zero loss
    assert loss < 1e-20


def test_parameter_count_matches_library(built):
    from myfvcomcode_amd.closure import NPARAM
    L = F.gpu_lib()
    assert L.fvmlp_nparams() == mlp.NPARAM == NPARAM


@pytest.mark.gpu
def test_closure_features_match_reference(built):
    m = F.Model.build(0, nx=21, ny=7, kb=6)
    dev = F.DeviceModel(m)
    print("case: channel 21x7 KB=6, features after 4 steps", flush=True)
    dev.run(4, graph=False)
    x, y = dev.closure_samples()
    import math
    ramp = math.tanh(4 / m.hdr.p.iramp)
    xr, yr = mlp.features(m, dev.get, ramp)
    assert x.shape == xr.shape
    np.testing.assert_allclose(x, xr, rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(y, yr, rtol=1e-6, atol=1e-30)


@pytest.mark.gpu
def test_training_gradient_matches_cpu(built):
    x, y = _samples(1500)
    t = F.Trainer(x, y, bmax=1024)
    p = _rand_params(seed=5)
    t.set_params(p)
    print("case: MLP gradient on 1000 random samples", flush=True)
    idx = np.arange(100, 1100, dtype=np.int32)
    loss, g = t.grad(idx)
    lr, gr = mlp.loss_grad(p.astype(np.float64), x[idx], y[idx])
    assert loss == pytest.approx(lr, rel=1e-4)
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    np.testing.assert_allclose(g, gr, rtol=2e-2, atol=1e-4 * np.abs(gr).max())


@pytest.mark.gpu
def test_mfma_inference_matches_fp32(built):
    x, y = _samples(3000)
    t = F.Trainer(x, y, bmax=1024)
    t.set_params(_rand_params(seed=7))
    print("case: MLP inference, BF16 MFMA vs FP32, 3000 samples", flush=True)
    k32 = t.infer(0, 3000, mfma=False)
    kmf = t.infer(0, 3000, mfma=True)
    kref = mlp.forward(t.params().astype(np.float64), x)
    np.testing.assert_allclose(k32, kref, rtol=1e-4)
    err = np.abs(np.log(kmf) - np.log(k32))
    assert err.max() < 0.1 and err.mean() < 0.02, (err.max(), err.mean())


def _channel_samples(steps=120, every=20):
    from myfvcomcode_amd.closure import collect
    return collect(0, steps, every)


@pytest.mark.gpu
def test_training_reduces_validation_loss(built):
    from myfvcomcode_amd.closure import train
    x, y = _channel_samples()
    print("case: train on %d channel samples" % len(x), flush=True)
    p, rep = train(x, y, steps=400, batch=4096, seed=0)
    print(rep, flush=True)
    assert rep["val_loss"] < 0.25 * rep["val_loss_init"]
    assert abs(rep["val_loss_mfma"] - rep["val_loss"]) < 0.2 * rep["val_loss"] + 0.02


@pytest.mark.gpu
def test_fallback_to_my25_for_unrealizable_columns(built):
    from myfvcomcode_amd.closure import init_params
    m = F.Model.build(0)
    p = init_params(0)
    p[-2:] = 30.0          # softplus(30) = 30 m^2/s, outside [1e-6, 1]: every column is rejected
    dev = F.DeviceModel(m)
    dev.set_closure(1, p)
    print("case: learned closure with unrealizable weights, 5 steps", flush=True)
    dev.run(5, graph=True)
    nf, nc = dev.closure_fallbacks()
    assert nf == nc == m.m
    km = dev.get("km").reshape(m.kb, -1)[1:-1, 1:]
    assert np.isfinite(km).all() and km.max() < 1.0


@pytest.mark.gpu
def test_learned_closure_gains_over_constant_mixing(built):
    from myfvcomcode_amd.closure import train, judge
    x, y = _channel_samples(240, 20)
    print("case: channel, train on %d samples, then 2 h A/B" % len(x), flush=True)
    p, rep = train(x, y, steps=800, batch=4096, seed=0)
    out = judge(p, case=0, hours=2.0)
    print(out, flush=True)
    L, C = out["learned"], out["constant"]
    assert L["current_rms"] < C["current_rms"]
    assert L["el_rms_m"] < C["el_rms_m"]
    assert L["fallback_columns_last_step"] < m_nodes(0)


def m_nodes(case):
    return F.Model.build(case).m
