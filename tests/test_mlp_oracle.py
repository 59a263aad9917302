"""Pins of the warm-start network oracle (oracle/mlp_oracle.py) by what the
mathematics fixes: finite differences of the loss, Adam's closed form at its
first step, and BF16 roundings known bit for bit."""
import numpy as np
import pytest

import mlp_oracle as mo


def small_net(rng, dims=(7, 16, 16, 3)):
    return [(rng.normal(0, 0.6, size=(dims[i + 1], dims[i])), rng.normal(0, 0.2, size=dims[i + 1]))
            for i in range(len(dims) - 1)]


def test_gradients_match_central_differences():
    rng = np.random.default_rng(11)
    params = small_net(rng)
    X = rng.normal(size=(9, 7))
    T = rng.normal(size=(9, 3))
    loss, grads = mo.loss_and_grads(params, X, T)
    h = 1e-6
    for l, (W, b) in enumerate(params):
        for arr, g in ((W, grads[l][0]), (b, grads[l][1])):
            flat = arr.reshape(-1)
            gflat = g.reshape(-1)
            for i in rng.choice(flat.size, size=min(12, flat.size), replace=False):
                old = flat[i]
                flat[i] = old + h
                lp = mo.loss_and_grads(params, X, T)[0]
                flat[i] = old - h
                lm = mo.loss_and_grads(params, X, T)[0]
                flat[i] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - gflat[i]) <= 1e-7 * max(1.0, abs(fd)), (l, i, fd, gflat[i])


def test_gelu_derivative_and_values():
    z = np.linspace(-6, 6, 121)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(z))) < 1e-8
    # GELU(0) = 0, GELU(z) -> z for large z, -> 0 for very negative z
    assert mo.gelu(0.0) == 0.0
    assert abs(mo.gelu(10.0) - 10.0) < 1e-12
    assert abs(mo.gelu(-10.0)) < 1e-12
    # GELU(1) = Phi(1) = 0.8413447460685429...
    assert abs(mo.gelu(1.0) - 0.8413447460685429) < 1e-15


def test_adam_first_step_closed_form():
    # at t = 1 the bias-corrected moments are g and g^2: the step is
    # lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    w = np.array([1.0, 1.0, 1.0, 1.0])
    w1, m, v = mo.adam_step(w, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    expect = w - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(w1, expect, rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bf16_rounding_known_values():
    ulp = 2.0 ** -7
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -(1 + 3 * 2 ** -8), 3.140625,
                     65280.0, 2.0 ** -130])
    out = mo.bf16_round(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                      # tie -> even
    assert out[2] == 1 + 2 * ulp              # tie -> even (mantissa 2)
    assert out[3] == 1 + ulp                  # above the tie
    assert out[4] == -(1 + 2 * ulp)
    assert out[5] == 3.140625                 # exactly representable
    assert out[6] == 65280.0
    assert out[7] == 2.0 ** -130              # subnormal power of two survives


def test_bf16_product_differs_from_fp64_by_rounding_only():
    rng = np.random.default_rng(3)
    params = small_net(rng, (32, 64, 4))
    X = rng.normal(size=(5, 32))
    Y64, _ = mo.forward(params, X)
    Y16, _ = mo.forward(params, X, bf16=True)
    rel = np.max(np.abs(Y16 - Y64)) / np.max(np.abs(Y64))
    assert 1e-5 < rel < 3e-2


def test_features_slots():
    total = np.array([[1e-3, -2e-4, 0.0]])
    prev = np.array([[1e-3, 2e-7, 0.0]])
    X = mo.features(total, prev, np.array([30.0]), np.array([[0.05, 0.02]]), 12, 3, 2)
    assert X[0, 0] == pytest.approx(np.log10(1 + 1e9))
    assert X[0, 1] == pytest.approx(-np.log10(1 + 2e8))
    assert X[0, 2] == 0.0
    assert X[0, 3] == pytest.approx(np.log(1e-3))
    assert X[0, 4] == pytest.approx(np.log(2e-7))
    assert X[0, 5] == pytest.approx(np.log(1e-300))   # a zero free ion is held finite
    assert X[0, 6] == 30.0
    assert list(X[0, 7:9]) == [0.05, 0.02]
    assert np.all(X[0, 9:] == 0.0)
