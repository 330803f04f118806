"""Pins of the CPU restatement of the learned samplers' network (oracle/mlp.py): the
gradient against central finite differences, the tanh-form GELU against the exact
x Phi(x), and bf16 rounding against exact bit patterns."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as M  # noqa: E402


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, M.NPARAM)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(1)
    X = rng.normal(size=(16, M.IN))
    T = rng.normal(size=(16, 2))
    loss, g = M.loss_grad(p, X, T, 1.0, 0.7)
    assert g.shape == (M.NPARAM,)
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False),
                          [M.P_B1, M.P_W2, M.P_B3 + 5, M.P_W4, M.P_B4, M.P_B4 + 1]])
    h = 1e-6
    for q in idx:
        e = np.zeros(M.NPARAM)
        e[q] = h
        fd = (M.loss_grad(p + e, X, T, 1.0, 0.7)[0] - M.loss_grad(p - e, X, T, 1.0, 0.7)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
    # the loss is the mean weighted squared error of the forward pass
    y = M.forward(p, X)
    assert loss == np.sum(np.array([1.0, 0.7]) * (y - T) ** 2) / 16


def test_gelu_tanh_form_tracks_exact_gelu():
    z = np.linspace(-6, 6, 2001)
    exact = np.array([0.5 * v * (1.0 + math.erf(v / math.sqrt(2.0))) for v in z])
    assert np.max(np.abs(M.gelu(z) - exact)) < 5e-4
    # derivative by differences
    h = 1e-6
    assert np.max(np.abs(M.gelu_grad(z) - (M.gelu(z + h) - M.gelu(z - h)) / (2 * h))) < 1e-8


def test_bf16_rounding_nearest_even():
    cases = {1.0: 1.0, 1.0 + 2 ** -8: 1.0, 1.0 + 3 * 2 ** -9: 1.0 + 2 ** -7, 1.0 + 2 ** -7 + 2 ** -8: 1.0 + 2 ** -6,
             -2.5: -2.5, 3.0e-3: 0.0030059814453125}
    for x, want in cases.items():
        assert float(M.bf16(np.float32(x))) == want, x


def test_bf16_forward_is_close_to_fp64_forward():
    rng = np.random.default_rng(5)
    p = random_params(2, 0.2)
    X = M.features(10 ** rng.uniform(4, 8, 500), rng.choice([7, 8, 18], 500), rng.random(500), rng.random(500))
    y64 = M.forward(p, X)
    y16 = M.forward_bf16(p, X)
    assert np.max(np.abs(y64 - y16)) < 0.05 * (1 + np.abs(y64).max())
