"""Pins of the network oracle (oracle/mlp_oracle.py) on the CPU: the
float64 backward pass against central finite differences, the bf16
rounding against hand-worked cases, and the device-rounded gradient against
the float64 one."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_cases():
    f = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5,
                  1.0 + 2 ** -8 + 2 ** -20], np.float32)
    # ties go to even; above the tie rounds up
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 1.0 + 2 ** -7],
                    np.float32)
    assert np.array_equal(mo.bf16(f), want)


def test_param_layout():
    assert mo.N_PARAMS == 12 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2 == 5122
    p = np.arange(mo.N_PARAMS, dtype=np.float64)
    assert np.array_equal(mo.pack(mo.unpack(p)), p)


def test_grad64_against_finite_differences():
    rng = np.random.default_rng(1)
    p = mo.init_params(3).astype(np.float64)
    z = rng.normal(size=(40, 12))
    t = rng.normal(size=(40, 2))
    g = mo.grad64(p, z, t, (1.0, 0.5))
    fd = np.empty_like(p)
    h = 1e-6
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mo.loss64(p + e, z, t, (1.0, 0.5)) - mo.loss64(p - e, z, t, (1.0, 0.5))) / (2 * h)
    assert np.linalg.norm(g - fd) <= 1e-6 * np.linalg.norm(fd)


def test_bf16_gradient_close_to_float64():
    rng = np.random.default_rng(2)
    p = mo.init_params(4)
    x = rng.normal(size=(256, 12)).astype(np.float32)
    t = rng.normal(size=(256, 2)).astype(np.float32)
    mean, std = np.zeros(12, np.float32), np.ones(12, np.float32)
    g16, _ = mo.grad_bf16(p, x, t, mean, std)
    g64 = mo.grad64(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert np.linalg.norm(g16 - g64) <= 2e-2 * np.linalg.norm(g64)
