"""Pins the CPU reference of the learned-stop MLP (oracle/mlp_ref.py) by
central finite differences, and its BF16 rounding by hand-made cases."""
import numpy as np

import mlp_ref


def test_bf16_rounding_cases():
    # exactly representable values are unchanged
    v = np.array([1.0, -2.5, 0.0, 3.0 * 2.0 ** 100], dtype=np.float32)
    assert np.array_equal(mlp_ref.bf16_round(v), v)
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1+2^-7: ties go to
    # the even mantissa (1.0); just above the midpoint rounds up
    half = np.float32(1.0 + 2.0 ** -8)
    above = np.nextafter(half, np.float32(2.0))
    assert mlp_ref.bf16_round(np.array([half]))[0] == np.float32(1.0)
    assert mlp_ref.bf16_round(np.array([above]))[0] == np.float32(1.0 + 2.0 ** -7)
    # 1 + 3*2^-8 is halfway between 1+2^-7 (odd) and 1+2^-6 (even)
    half2 = np.float32(1.0 + 3 * 2.0 ** -8)
    assert mlp_ref.bf16_round(np.array([half2]))[0] == np.float32(1.0 + 2.0 ** -6)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(size=mlp_ref.P) * 0.3
    x = rng.normal(size=(12, 16))
    y = (rng.random(12) > 0.5).astype(np.float64)
    w = rng.random(12) + 0.5
    _, g, _ = mlp_ref.loss_and_grad(p, x, y, w, round_bf16=False)
    eps = 1e-6
    idx = rng.choice(mlp_ref.P, size=60, replace=False)
    for i in idx:
        pp = p.copy(); pp[i] += eps
        pm = p.copy(); pm[i] -= eps
        lp, _, _ = mlp_ref.loss_and_grad(pp, x, y, w, round_bf16=False)
        lm, _, _ = mlp_ref.loss_and_grad(pm, x, y, w, round_bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_zero_weight_samples_do_not_contribute():
    rng = np.random.default_rng(1)
    p = rng.normal(size=mlp_ref.P) * 0.3
    x = rng.normal(size=(8, 16))
    y = np.ones(8)
    w = np.array([1, 0, 1, 0, 1, 0, 1, 0], dtype=np.float64)
    l1, g1, _ = mlp_ref.loss_and_grad(p, x, y, w)
    l2, g2, _ = mlp_ref.loss_and_grad(p, x[::2], y[::2], w[::2])
    assert np.isclose(l1, l2) and np.allclose(g1, g2)
