"""The learned-Q network on the matrix cores against its CPU oracle
(oracle/mlp_oracle.py, device-rounded form): forward pass, gradient of the
loss, and an Adam training run that fits a known function."""
import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def make(lulesh_gpu, seed=0):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p[mo.N_PARAMS - 2:] = [0.3, -0.2]          # nonzero output biases
    mean = rng.normal(size=12).astype(np.float32)
    std = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    return lulesh_gpu.Mlp(p, mean, std), p, mean, std


@pytest.mark.parametrize("n", [1, 63, 64, 200, 1000])
def test_forward_matches_oracle(lulesh_gpu, n):
    net, p, mean, std = make(lulesh_gpu, n)
    x = np.random.default_rng(10 + n).normal(size=(n, 12)).astype(np.float32) * std + mean
    y = net.eval(x)
    want, _ = mo.forward_bf16(p, x, mean, std)
    scale = np.max(np.abs(want)) + 1e-6
    # same bf16 rounding points; float32 sums in another order may flip a
    # bf16 activation by one unit now and then
    assert np.max(np.abs(y - want)) <= 2e-2 * scale
    assert np.median(np.abs(y - want)) <= 1e-4 * scale


@pytest.mark.parametrize("n", [64, 333, 2048])
def test_gradient_matches_oracle(lulesh_gpu, n):
    net, p, mean, std = make(lulesh_gpu, 7)
    rng = np.random.default_rng(n)
    x = (rng.normal(size=(n, 12)) * std + mean).astype(np.float32)
    t = rng.normal(size=(n, 2)).astype(np.float32)
    g, loss = net.gradient(x, t, (1.0, 0.25))
    g_ref, loss_ref = mo.grad_bf16(p, x, t, mean, std, (1.0, 0.25))
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    ref = mo.unpack(g_ref.astype(np.float64))
    got = mo.unpack(g.astype(np.float64))
    for k in ref:
        err = np.linalg.norm(got[k] - ref[k])
        assert err <= 2e-2 * np.linalg.norm(ref[k]) + 1e-7, (k, err, np.linalg.norm(ref[k]))
    # and close to the exact float64 gradient
    z = (x - mean) / std
    g64 = mo.grad64(p.astype(np.float64), z.astype(np.float64), t.astype(np.float64),
                    (1.0, 0.25))
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)


def test_training_fits_a_function(lulesh_gpu):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(8192, 12)).astype(np.float32)
    t = np.stack([np.maximum(x[:, 0] + 0.5 * x[:, 1], 0.0),
                  0.5 * x[:, 2] * x[:, 3]], axis=1).astype(np.float32)
    p = mo.init_params(11)
    net = lulesh_gpu.Mlp(p, np.zeros(12, np.float32), np.full(12, 0.5, np.float32))
    _, loss0 = net.gradient(x, t)
    loss = net.train(x, t, epochs=30, batch=512, lr=3e-3, seed=1)
    _, loss1 = net.gradient(x, t)
    assert loss1 < 0.1 * loss0, (loss0, loss, loss1)
    # deterministic: the same run gives the same weights
    net2 = lulesh_gpu.Mlp(p, np.zeros(12, np.float32), np.full(12, 0.5, np.float32))
    net2.train(x, t, epochs=30, batch=512, lr=3e-3, seed=1)
    assert np.array_equal(net.params(), net2.params())
