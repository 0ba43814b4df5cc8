"""numpy restatement of the learned closure's per-point stage (les_model_type 6):
the 13 features of a GY point, the training-envelope test and the GY
coefficients the step uses, from the type-5 quantities the oracle captures
(oracle_les_capture) and the network of oracle/mlp.py.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker of the device's
k_closure and k_amd<..., 6>.  The learned closure has no counterpart in the
reference; it replaces the AMD coefficient of les_chan / les_chan_th type 5
(for/les.f90:571-797, 1201-1348) by the network's prediction.

Features (float32 arithmetic), from the scaled strain S = [[a1 a4 a5][a4 a2 a6]
[a5 a6 a3]], rotation O = [[0 o4 o5][-o4 0 o6][-o5 -o6 0]], P = O O^T, the scaled
scalar gradient b, s2 = S:S, w2 = O:O, e = s2 + w2, with n/d taken as 0 where
d <= 0:
  0 (w2 - s2)/e            1 tr S^3/s2^1.5          2 tr(S P)/(s2^.5 w2)
  3 -(tr S^3 + tr SP)/(s2^.5 e)   4 tr(S^2 P)/(s2 w2)   5 b_y^2/|b|^2
  6 b.S.b/(|b|^2 s2^.5)   7 b.S^2.b/(|b|^2 s2)   8 b.P.b/(|b|^2 w2)
  9 tanh(g_y th_y / s2)   10 tanh(|g_y| |grad th| / s2)
 11 log(1 + D2 sqrt(2 s2)/nu)   12 log(delta_y^2/(delta_x delta_z))
D2 = 3/(1/delta_x^2 + 1/delta_y^2 + 1/delta_z^2), delta_x = dx beta, delta_z = dz beta,
delta_y = 2 dy(j) (the type-5 filter widths at GY, for/les.f90:756-780).
"""
from __future__ import annotations

import numpy as np

from . import mlp

BETA_SGS = 3.0   # for/les.f90:43
C_AMD = 0.2887   # for/les.f90:25
CAP_A1, CAP_O4, CAP_V, CAP_BX, CAP_TY, CAP_VT = 0, 6, 9, 10, 13, 14


def features(cap, cfg, dy):
    """X [ny2][nz][nx][16] float32 features at the GY points of rows 2..Nyp (zero
    elsewhere) and s2, from the capture array of an oracle LES substep."""
    f32 = np.float32
    c = cap.astype(f32)
    a1, a2, a3, a4, a5, a6 = (c[..., CAP_A1 + k] for k in range(6))
    o4, o5, o6 = (c[..., CAP_O4 + k] for k in range(3))
    bx, by, bz = (c[..., CAP_BX + k] for k in range(3))
    ty = c[..., CAP_TY]
    dxs, dzs = cfg.lx / cfg.nx * BETA_SGS, cfg.lz / cfg.nz * BETA_SGS
    s2 = a1 * a1 + a2 * a2 + a3 * a3 + f32(2) * (a4 * a4 + a5 * a5 + a6 * a6)
    w2 = f32(2) * (o4 * o4 + o5 * o5 + o6 * o6)
    e = s2 + w2
    S = np.stack([np.stack([a1, a4, a5], -1), np.stack([a4, a2, a6], -1), np.stack([a5, a6, a3], -1)], -2)
    z = np.zeros_like(a1)
    O = np.stack([np.stack([z, o4, o5], -1), np.stack([-o4, z, o6], -1), np.stack([-o5, -o6, z], -1)], -2)
    S2 = np.einsum("...im,...mk->...ik", S, S)
    P = np.einsum("...im,...km->...ik", O, O)
    trS3 = np.einsum("...ik,...ki->...", S2, S)
    trSP = np.einsum("...ik,...ki->...", S, P)
    trS2P = np.einsum("...ik,...ki->...", S2, P)
    b = np.stack([bx, by, bz], -1)
    b2 = bx * bx + by * by + bz * bz
    bSb = np.einsum("...i,...ik,...k->...", b, S, b)
    bS2b = np.einsum("...i,...ik,...k->...", b, S2, b)
    bPb = np.einsum("...i,...ik,...k->...", b, P, b)
    rs = np.sqrt(s2)

    def q(n, d):
        out = np.zeros_like(n)
        np.divide(n, d, out=out, where=d > 0)
        return out

    tx, tz = bx / f32(dxs), bz / f32(dzs)
    gth = np.sqrt(tx * tx + ty * ty + tz * tz)
    gyb = f32(cfg.grav_y)
    ny, n2 = cfg.ny, cfg.ny2
    dyl = (dy * 2.0).astype(f32)[:, None, None]
    with np.errstate(divide="ignore"):  # dy(0) = 0: row 0 is not a GY point
        D2 = (3.0 / (1.0 / dxs ** 2 + 1.0 / (2.0 * dy) ** 2 + 1.0 / dzs ** 2)).astype(f32)[:, None, None]
    X = np.zeros(a1.shape + (16,), f32)
    X[..., 0] = q(w2 - s2, e)
    X[..., 1] = q(trS3, s2 * rs)
    X[..., 2] = q(trSP, rs * w2)
    X[..., 3] = q(-(trS3 + trSP), rs * e)
    X[..., 4] = q(trS2P, s2 * w2)
    X[..., 5] = q(by * by, b2)
    X[..., 6] = q(bSb, b2 * rs)
    X[..., 7] = q(bS2b, b2 * s2)
    X[..., 8] = q(bPb, b2 * w2)
    X[..., 9] = np.tanh(q(gyb * ty, s2))
    X[..., 10] = np.tanh(q(np.abs(gyb) * gth, s2))
    X[..., 11] = np.log1p(D2 * np.sqrt(f32(2) * s2) / f32(cfg.nu))
    with np.errstate(divide="ignore"):
        X[..., 12] = np.broadcast_to(np.log(dyl * dyl / f32(dxs * dzs)), a1.shape)
    rows = np.zeros(n2, bool)
    rows[2:ny + 1] = True
    X[~rows] = 0.0
    s2 = np.where(rows[:, None, None], s2, f32(0))
    return X, s2


def coefficients(cap, cfg, dy, params, norm):
    """The GY coefficients of les_model_type 6 ([ny2][nz][nx] each): where every
    feature lies in the envelope [lo, hi] of norm and s2 > 0, sqrt(2 s2) max(0, y)
    of the network (y0 velocity, y1 scalar); elsewhere the type-5 coefficient with
    its sign bit set.  Also returns the inside mask."""
    X, s2 = features(cap, cfg, dy)
    norm = np.asarray(norm, np.float32)
    lo, hi = norm[32:32 + mlp.NIN], norm[48:48 + mlp.NIN]
    F = X[..., :mlp.NIN]
    inside = (s2 > 0) & np.all((F >= lo) & (F <= hi), axis=-1)
    y, _ = mlp.forward(params, norm, X.reshape(-1, 16), emulate=True)
    y = y.astype(np.float32).reshape(X.shape[:-1] + (2,))
    r2 = np.sqrt(np.float32(2) * s2)
    cgy = np.where(inside, (r2 * np.maximum(np.float32(0), y[..., 0])).astype(np.float64), -cap[..., CAP_V])
    kgy = np.where(inside, (r2 * np.maximum(np.float32(0), y[..., 1])).astype(np.float64), -cap[..., CAP_VT])
    return cgy, kgy, inside
