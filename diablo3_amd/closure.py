"""The learned LES closure network on the device (include/dgpu_closure.h).

``Mlp`` wraps a dgpu_mlp handle: an MLP 13 -> 64 -> 64 -> 2 (GELU) whose
forward pass, backward pass and Adam run as BF16 MFMA kernels with FP32
accumulation.  Training data are the features of dgpu_les_features (invariants
of the resolved gradients plus the filter width, evaluated on box-filtered DNS)
and targets from the filtered DNS's own subgrid stresses (diablo3_amd.sgs_data).
There is no CPU fallback: without the HIP library construction raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .solver import load

NFEAT, NFEAT_PAD, NOUT = 13, 16, 2
NPARAM = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 16 + 16
NNORM = 64
C_AMD = 0.2887  # the type-5 constant (for/les.f90:25)


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc):
    if rc:
        raise RuntimeError(f"libdiablo_gpu error {rc}: {load().dgpu_last_error().decode()}")


def make_norm(X: np.ndarray, envelope_quantile: float = 0.0, max_rows: int = 10_000_000) -> np.ndarray:
    """Normaliser from training features X [n][16]: shift = mean, scale = 1/std
    on the 13 live features, and the envelope [lo, hi] = the [q, 1-q] quantiles
    (q = 0: min and max, over all rows).  Mean, std and quantiles of more than
    max_rows rows come from an evenly strided subset of max_rows rows."""
    step = max(1, X.shape[0] // max_rows)
    Xs = np.asarray(X[::step, :NFEAT], dtype=np.float64)
    n = np.zeros(NNORM, dtype=np.float32)
    sd = Xs.std(axis=0)
    n[:NFEAT] = Xs.mean(axis=0)
    n[16:16 + NFEAT] = np.where(sd > 0, 1.0 / np.where(sd > 0, sd, 1.0), 0.0)
    if envelope_quantile > 0:
        lo = np.quantile(Xs, envelope_quantile, axis=0)
        hi = np.quantile(Xs, 1.0 - envelope_quantile, axis=0)
    else:
        lo, hi = X[:, :NFEAT].min(axis=0), X[:, :NFEAT].max(axis=0)
    n[32:48] = -np.inf
    n[48:64] = np.inf
    n[32:32 + NFEAT] = lo
    n[48:48 + NFEAT] = hi
    return n


DEFAULT_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "closure_front.npz")


def load_params(path: Optional[str] = None):
    """(params, norm, source) from a .npz written by train_closure, the committed
    front closure, or (neither present) a seeded network with an unbounded
    envelope."""
    path = path or (DEFAULT_FILE if os.path.exists(DEFAULT_FILE) else None)
    if path:
        z = np.load(path, allow_pickle=False)
        return z["params"].astype(np.float32), z["norm"].astype(np.float32), os.path.basename(path)
    m = Mlp(seed=0)
    n = np.zeros(NNORM, dtype=np.float32)
    n[16:16 + NFEAT] = 1.0
    n[32:48] = -np.inf
    n[48:64] = np.inf
    return m.params, n, "seeded network (untrained)"


class Mlp:
    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 0):
        L = load()
        h = ctypes.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        _check(L.dgpu_mlp_create(None if p is None else _fp(p), ctypes.c_uint64(seed), ctypes.byref(h)))
        self.h = h
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            load().dgpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(NPARAM, dtype=np.float32)
        _check(load().dgpu_mlp_get_params(self.h, _fp(out)))
        return out

    @params.setter
    def params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAM
        _check(load().dgpu_mlp_set_params(self.h, _fp(p)))

    def set_norm(self, norm: np.ndarray):
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        assert norm.size == NNORM
        _check(load().dgpu_mlp_set_norm(self.h, _fp(norm)))
        self.norm = norm.copy()

    def set_data(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        assert X.ndim == 2 and X.shape[1] == NFEAT_PAD and T.shape == (X.shape[0], NOUT)
        _check(load().dgpu_mlp_set_data(self.h, _fp(X), _fp(T), X.shape[0]))
        self.n = X.shape[0]

    def grad(self, n: int, loss_w: Sequence[float] = (1.0, 1.0)):
        g = np.zeros(NPARAM, dtype=np.float32)
        lw = np.ascontiguousarray(loss_w, dtype=np.float32)
        loss = ctypes.c_float()
        _check(load().dgpu_mlp_grad(self.h, n, _fp(lw), _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def adam_step(self, grad: np.ndarray, lr: float):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        _check(load().dgpu_mlp_adam_step(self.h, _fp(g), ctypes.c_float(lr)))

    def train(self, epochs: int, batch: int, lr: float = 1e-3, seed: int = 0,
              loss_w: Sequence[float] = (1.0, 1.0)) -> np.ndarray:
        out = np.zeros(epochs, dtype=np.float32)
        lw = np.ascontiguousarray(loss_w, dtype=np.float32)
        _check(load().dgpu_mlp_train(self.h, epochs, batch, ctypes.c_float(lr), ctypes.c_uint64(seed),
                                     _fp(lw), _fp(out)))
        return out

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], NOUT), dtype=np.float32)
        _check(load().dgpu_mlp_predict(self.h, _fp(X), X.shape[0], _fp(Y)))
        return Y
