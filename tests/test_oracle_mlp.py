"""Pins the binary64 MLP oracle (oracle/mlp_oracle.c): GELU values, the
gradient against central finite differences, a closed-form gradient of a
network reduced to a linear map, and the size and sign of one Adam step."""
import math

import numpy as np
import pytest

from oracle import oracle as O

NP = 5385
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 576, 640, 4736, 4800, 5376


def test_param_count():
    assert O.mlp_param_count() == NP == 9 * 64 + 64 + 64 * 64 + 64 + 64 * 9 + 9


def test_gelu_values():
    L = O.lib()
    assert L.mlp_gelu(0.0) == 0.0
    assert abs(L.mlp_gelu(1.0) - 0.8413447460685429) < 1e-15  # 1 * Phi(1)
    assert abs(L.mlp_gelu(-1.0) + 0.15865525393145707) < 1e-15
    assert abs(L.mlp_gelu(8.0) - 8.0) < 1e-12 and abs(L.mlp_gelu(-8.0)) < 1e-12
    assert L.mlp_gelu_grad(0.0) == 0.5
    for x in (-2.3, -0.4, 0.1, 1.7):
        h = 1e-6
        fd = (L.mlp_gelu(x + h) - L.mlp_gelu(x - h)) / (2 * h)
        assert abs(fd - L.mlp_gelu_grad(x)) < 1e-9


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(NP) * 0.2
    x = rng.standard_normal((7, 9))
    t = rng.standard_normal((7, 9))
    loss, g = O.mlp_loss_grad(p, x, t)
    idx = np.concatenate([rng.choice(np.arange(s, e), 6, replace=False) for s, e in
                          [(OFF_W1, OFF_B1), (OFF_B1, OFF_W2), (OFF_W2, OFF_B2), (OFF_B2, OFF_W3),
                           (OFF_W3, OFF_B3), (OFF_B3, NP)]])
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.mlp_loss_grad(pp, x, t)[0] - O.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])), k


def test_output_layer_gradient_closed_form():
    """With W1 = W2 = 0 and b1, b2 set, the hidden activations are the
    constants gelu(b1), gelu(gelu(b1)) W2 + b2 = gelu(b2); then
    y = W3 h + b3 and dL/db3 = 2/(9B) sum_b (y_b - t_b)."""
    rng = np.random.default_rng(4)
    p = np.zeros(NP)
    p[OFF_B1:OFF_W2] = rng.standard_normal(64)
    p[OFF_B2:OFF_W3] = rng.standard_normal(64)
    p[OFF_W3:OFF_B3] = rng.standard_normal(9 * 64) * 0.1
    p[OFF_B3:] = rng.standard_normal(9)
    gelu = np.vectorize(O.lib().mlp_gelu)
    h2 = gelu(p[OFF_B2:OFF_W3])
    y = p[OFF_W3:OFF_B3].reshape(9, 64) @ h2 + p[OFF_B3:]
    x = rng.standard_normal((5, 9))
    t = rng.standard_normal((5, 9))
    np.testing.assert_allclose(O.mlp_forward(p, x), np.broadcast_to(y, (5, 9)), rtol=1e-13)
    loss, g = O.mlp_loss_grad(p, x, t)
    r = y[None, :] - t
    np.testing.assert_allclose(loss, np.mean(r * r), rtol=1e-13)
    np.testing.assert_allclose(g[OFF_B3:], 2.0 / 45.0 * r.sum(axis=0), rtol=1e-12)
    np.testing.assert_allclose(g[OFF_W3:OFF_B3].reshape(9, 64), 2.0 / 45.0 * np.outer(r.sum(axis=0), h2), rtol=1e-12)
    assert np.all(g[OFF_W1:OFF_B1] == 0.0)  # x only reaches the loss through W1 = 0 ... W2 = 0


def test_first_adam_step_is_lr_times_sign():
    """With bias correction the first update is lr * g / (|g| + eps)."""
    rng = np.random.default_rng(5)
    p = rng.standard_normal(NP)
    g = rng.standard_normal(NP)
    m = np.zeros(NP)
    v = np.zeros(NP)
    p0 = p.copy()
    O.mlp_adam(p, g, m, v, 1, 1e-3)
    np.testing.assert_allclose(p0 - p, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-10)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-14)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-14)
