"""Pins of the network oracle (oracle/mlp.py) by what the mathematics fixes:
central finite differences of the loss for the backward pass, the closed form of
Adam under a constant gradient, BF16 round-to-nearest-even on known values, the
GELU closed values and derivative."""
import numpy as np

from oracle import mlp


def rand_params(rng, scale=0.3):
    p = rng.uniform(-scale, scale, mlp.NP)
    p[~mlp.live_mask()] = 0.0
    return p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -1 - 3 * 2 ** -8, 3.0, 0.0],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -1 - 2 ** -6, 3.0, 0.0], dtype=np.float32)
    assert np.array_equal(mlp.bf16(x), want)


def test_gelu_values_and_derivative():
    from scipy.special import erf
    assert mlp.gelu(0.0) == 0.0
    assert abs(mlp.gelu(8.0) - 8.0) < 1e-12 and abs(mlp.gelu(-8.0)) < 1e-12
    # the tanh form stays within 2e-3 of the erf definition x Phi(x)
    x = np.linspace(-6, 6, 241)
    assert np.max(np.abs(mlp.gelu(x) - 0.5 * x * (1 + erf(x / np.sqrt(2))))) < 2e-3
    x = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.dgelu(x))) < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rand_params(rng)
    X = np.zeros((37, 16), np.float32)
    X[:, :13] = rng.standard_normal((37, 13))
    T = rng.standard_normal((37, 2))
    norm = mlp.default_norm()
    lw = (1.0, 0.5)
    _, g = mlp.loss_grad(p, norm, X, T, lw, emulate=False)
    live = np.flatnonzero(mlp.live_mask())
    pick = np.concatenate([rng.choice(live, 300, replace=False), np.arange(mlp.OB1, mlp.OW2),
                           np.arange(mlp.OB3, mlp.OB3 + 2)])
    h = 1e-6
    for i in pick:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, norm, X, T, lw, emulate=False)[0] -
              mlp.loss_grad(pm, norm, X, T, lw, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])
    assert np.all(g[~mlp.live_mask()] == 0)


def test_emulated_gradient_close_to_exact():
    rng = np.random.default_rng(4)
    p = rand_params(rng)
    X = np.zeros((200, 16), np.float32)
    X[:, :13] = rng.standard_normal((200, 13))
    T = rng.standard_normal((200, 2))
    norm = mlp.default_norm()
    le, ge = mlp.loss_grad(p, norm, X, T, emulate=True)
    lx, gx = mlp.loss_grad(p, norm, X, T, emulate=False)
    assert abs(le - lx) < 2e-2 * lx
    assert np.linalg.norm(ge - gx) < 3e-2 * np.linalg.norm(gx)


def test_adam_constant_gradient_closed_form():
    """With a constant gradient g the bias-corrected moments are g and g^2 at
    every step, so each step moves every live parameter by lr g/(|g| + eps)."""
    rng = np.random.default_rng(5)
    p0 = rand_params(rng)
    g = rng.standard_normal(mlp.NP)
    g[~mlp.live_mask()] = 0.0
    p, m, v = p0.copy(), np.zeros(mlp.NP), np.zeros(mlp.NP)
    lr = 1e-3
    for t in range(1, 6):
        p, m, v = mlp.adam(p, m, v, g, t, lr)
        want = p0 - t * lr * g / (np.abs(g) + 1e-8)
        want[~mlp.live_mask()] = 0.0
        assert np.max(np.abs(p - want)) < 1e-12


def test_forward_constant_output_layer():
    rng = np.random.default_rng(6)
    p = rand_params(rng)
    p[mlp.OW3:mlp.OB3] = 0.0
    p[mlp.OB3:mlp.OB3 + 2] = [0.25, -0.5]
    X = np.zeros((5, 16), np.float32)
    X[:, :13] = rng.standard_normal((5, 13))
    y, _ = mlp.forward(p, mlp.default_norm(), X)
    assert np.array_equal(y, np.tile([0.25, -0.5], (5, 1)))
