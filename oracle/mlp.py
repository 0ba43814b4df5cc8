"""numpy restatement of the learned closure network (MLP 13 -> 64 -> 64 -> 2,
GELU in its tanh form) and of its training step: forward pass, backward pass and Adam, with the
BF16 rounding points of the device kernels reproduced (``emulate=True``) or left
out (``emulate=False``, plain float64 backpropagation).

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker of the device's
network kernels.  The network has no counterpart in the reference; it is a
further LES model beside les_chan's (for/les.f90:3-1435).

Parameter block layout (float, NP entries): W1 [16][64], b1 [64], W2 [64][64],
b2 [64], W3 [64][16], b3 [16] (W stored [in][out]; rows 13..15 of W1 and
columns 2..15 of W3, b3 zero).  Normaliser: shift[16], scale[16], lo[16], hi[16].
Loss: L = (1/n) sum_s sum_o w_o (y_so - t_so)^2.
"""
from __future__ import annotations

import numpy as np

NIN, NINP, NH, NOUT, NOUTP = 13, 16, 64, 2, 16
OW1 = 0
OB1 = OW1 + NINP * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * NOUTP
NP = OB3 + NOUTP
NNORM = 4 * NINP


def bf16(x):
    """float32 -> nearest-even bfloat16, returned as float32 (finite inputs)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OW1:OB1].reshape(NINP, NH)
    b1 = p[OB1:OW2]
    W2 = p[OW2:OB2].reshape(NH, NH)
    b2 = p[OB2:OW3]
    W3 = p[OW3:OB3].reshape(NH, NOUTP)
    b3 = p[OB3:NP]
    return W1, b1, W2, b2, W3, b3


def live_mask():
    m = np.ones(NP, dtype=bool)
    m[OW1:OB1] = (np.arange(NINP * NH) // NH) < NIN
    m[OW3:OB3] = (np.arange(NH * NOUTP) % NOUTP) < NOUT
    m[OB3:NP] = np.arange(NOUTP) < NOUT
    return m


K0, K1 = np.sqrt(2.0 / np.pi), 0.044715


def gelu(x):
    """GELU in its tanh form (Hendrycks & Gimpel, 2016): x (1 + tanh u)/2,
    u = sqrt(2/pi) (x + 0.044715 x^3)."""
    return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def dgelu(x):
    t = np.tanh(K0 * (x + K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


def default_norm():
    n = np.zeros(NNORM, dtype=np.float32)
    n[NINP:NINP + NIN] = 1.0
    n[2 * NINP:3 * NINP] = -np.inf
    n[3 * NINP:] = np.inf
    return n


def normalise(norm, X):
    norm = np.asarray(norm, dtype=np.float32)
    X = np.asarray(X, dtype=np.float32)
    return (X - norm[:NINP]) * norm[NINP:2 * NINP]


def forward(p, norm, X, emulate=True):
    """Outputs y [n][2] and the activations the backward pass needs."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(normalise(norm, X)).astype(np.float64)
    if emulate:
        w1, w2, w3 = (bf16(W.astype(np.float32)).astype(np.float64) for W in (W1, W2, W3))
    else:
        w1, w2, w3 = W1, W2, W3
    z1 = x @ w1 + b1
    h1 = r(gelu(z1).astype(np.float32) if emulate else gelu(z1)).astype(np.float64)
    z2 = h1 @ w2 + b2
    h2 = r(gelu(z2).astype(np.float32) if emulate else gelu(z2)).astype(np.float64)
    y = h2 @ w3 + b3
    return y[:, :NOUT], dict(x=x, w2=w2, w3=w3, z1=z1, h1=h1, z2=z2, h2=h2)


def loss_grad(p, norm, X, T, loss_w=(1.0, 1.0), emulate=True):
    """Loss and gradient (NP entries) of the mean weighted squared error.

    emulate=True rounds where the device kernels round: the normalised inputs,
    weights and activations (forward), dL/dy and dL/dz2 (the operands of the
    layer-3 and layer-2 products) and dL/dz1 (the operand of the layer-1 weight
    product) to BF16."""
    n = X.shape[0]
    T = np.asarray(T, dtype=np.float64)
    y, a = forward(p, norm, X, emulate)
    wo = np.asarray(loss_w, dtype=np.float64)
    e = y - T
    loss = float(np.sum(wo * e * e) / n)
    r = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if emulate else (lambda v: v)
    dy = np.zeros((n, NOUTP))
    dy[:, :NOUT] = 2.0 * wo * e / n
    if emulate:
        dy = dy.astype(np.float32).astype(np.float64)
    dyb = r(dy)
    g = np.zeros(NP)
    gW3 = a["h2"].T @ dyb
    gb3 = dy.sum(axis=0)
    dz2 = (dyb @ a["w3"].T) * dgelu(a["z2"])
    dz2b = r(dz2)
    gW2 = a["h1"].T @ dz2b
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2b @ a["w2"].T) * dgelu(a["z1"])
    gW1 = a["x"].T @ r(dz1)
    gb1 = dz1.sum(axis=0)
    g[OW1:OB1] = gW1.ravel()
    g[OB1:OW2] = gb1
    g[OW2:OB2] = gW2.ravel()
    g[OB2:OW3] = gb2
    g[OW3:OB3] = gW3.ravel()
    g[OB3:NP] = gb3
    g[~live_mask()] = 0.0
    return loss, g


def adam(p, m, v, grad, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, 2015) at step count t (>= 1); padded entries stay zero."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    p = p - lr * mh / (np.sqrt(vh) + eps)
    live = live_mask()
    p = np.where(live, p, 0.0)
    return p, np.where(live, m, 0.0), np.where(live, v, 0.0)
