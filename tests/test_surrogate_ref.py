"""Pins the CPU reference of the cooling-surrogate MLP (oracle/mlp_ref.py):
its backward pass against central finite differences of its own FP64
forward pass, its bf16 rounding against the IEEE definition, and its Adam
step against the closed form of the first step."""
import numpy as np

import mlp_ref as M


def test_parameter_count():
    assert M.NPARAM == 8770
    assert M.init_params(0).shape == (8770,)


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even (1.0)
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    # just above halfway rounds up
    assert M.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(1).normal(size=1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-6))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = M.init_params(5).astype(np.float64) + rng.normal(scale=0.05, size=M.NPARAM)
    x = rng.normal(size=(37, 4))
    t = rng.normal(size=(37, 2))
    L, g = M.loss_and_grad(theta, x, t, w=(1.0, 0.7), round_bf16=False)
    idx = rng.choice(M.NPARAM, size=60, replace=False)
    idx = np.concatenate([idx, [0, 255, 256, 8767, 8768, 8769]])   # first/last weights and biases
    h = 1e-6
    for k in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        Lp = M.loss_and_grad(tp, x, t, w=(1.0, 0.7), round_bf16=False)[0]
        Lm = M.loss_and_grad(tm, x, t, w=(1.0, 0.7), round_bf16=False)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.allclose(M.dsilu(z), fd, atol=1e-9)


def test_adam_first_step_is_signed_lr():
    theta = np.zeros(5)
    g = np.array([1e-3, -2.0, 0.5, -1e-6, 3.0])
    t1, m, v = M.adam(theta, g, np.zeros(5), np.zeros(5), 1, 0.01, eps=0.0)
    assert np.allclose(t1, -0.01 * np.sign(g))
