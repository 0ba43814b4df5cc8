"""The closure network's BF16 MFMA kernels against the numpy oracle
(oracle/mlp.py, with the kernels' BF16 rounding points), through the C ABI.

Bars: predictions within 2e-3 of the output scale (a BF16 rounding of an
activation may fall the other way when the FP32 sums differ in the last bit);
the full-batch gradient within 5e-3 relative L2 and the loss within 1e-5; Adam
steps within 1e-6 relative; training is deterministic (same seed, same
parameters bit for bit) and lowers the loss of a smooth regression."""
import numpy as np
import pytest

from diablo3_amd.closure import Mlp
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def setup(n, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, om.NP).astype(np.float32)
    p[~om.live_mask()] = 0
    X = np.zeros((n, 16), np.float32)
    X[:, :13] = rng.standard_normal((n, 13)) * 2 + 0.5
    norm = om.default_norm()
    norm[:13] = 0.5
    norm[16:29] = 0.5
    T = rng.standard_normal((n, 2)).astype(np.float32)
    return p, X, T, norm


@pytest.mark.parametrize("n", [16, 1000, 4099])
def test_predict_matches_oracle(n):
    p, X, _, norm = setup(n, 1)
    m = Mlp(p)
    m.set_norm(norm)
    y = m.predict(X)
    yo, _ = om.forward(p, norm, X, emulate=True)
    err = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    print(f"predict n={n}: max err {err:.2e}", flush=True)
    assert err <= 2e-3


@pytest.mark.parametrize("n", [16, 999, 20000])
def test_gradient_matches_oracle(n):
    p, X, T, norm = setup(n, 2)
    m = Mlp(p)
    m.set_norm(norm)
    m.set_data(X, T)
    lw = (1.0, 0.5)
    loss, g = m.grad(n, lw)
    lo, go = om.loss_grad(p, norm, X, T, lw, emulate=True)
    _, gx = om.loss_grad(p, norm, X, T, lw, emulate=False)
    e = np.linalg.norm(g - go) / np.linalg.norm(go)
    ex = np.linalg.norm(g - gx) / np.linalg.norm(gx)
    print(f"grad n={n}: vs emulated {e:.2e}, vs exact {ex:.2e}, loss {loss:.6f} / {lo:.6f}", flush=True)
    assert e <= 5e-3
    assert abs(loss - lo) <= 1e-5 * lo
    assert np.all(g[~om.live_mask()] == 0)


def test_adam_matches_oracle():
    p, _, _, _ = setup(16, 3)
    m = Mlp(p)
    rng = np.random.default_rng(7)
    pp, mm, vv = p.astype(np.float64), np.zeros(om.NP), np.zeros(om.NP)
    for t in range(1, 4):
        g = rng.standard_normal(om.NP).astype(np.float32)
        m.adam_step(g, 1e-2)
        pp, mm, vv = om.adam(pp, mm, vv, g.astype(np.float64), t, 1e-2)
    got = m.params
    err = np.max(np.abs(got - pp)) / np.max(np.abs(pp))
    print(f"adam x3: max err {err:.2e}", flush=True)
    assert err <= 1e-6


def test_training_deterministic_and_learns():
    rng = np.random.default_rng(9)
    n = 8192
    X = np.zeros((n, 16), np.float32)
    X[:, :13] = rng.uniform(-1, 1, (n, 13))
    T = np.stack([0.5 * np.tanh(2 * X[:, 0]) + 0.3 * X[:, 1] ** 2,
                  0.2 * np.sin(3 * X[:, 2]) * X[:, 3]], axis=1).astype(np.float32)
    runs = []
    for _ in range(2):
        m = Mlp(seed=5)
        m.set_data(X, T)
        loss0, _ = m.grad(n)
        hist = m.train(epochs=30, batch=256, lr=3e-3, seed=11)
        loss1, _ = m.grad(n)
        runs.append((m.params, hist, loss0, loss1))
    print("epoch losses", runs[0][1][[0, 9, 19, 29]], "full-batch", runs[0][2], "->", runs[0][3], flush=True)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][3] < 0.05 * runs[0][2]
