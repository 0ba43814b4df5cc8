"""The learned closure end to end on a small front: DNS on the device, box
filter onto an LES grid (4x in x and z, about 2x in y), device features and
least-squares targets, training by the device's kernels, and the gain over the
method without the network.

Gain (a priori, on snapshots of a held-out DNS seed): the relative mean-square
error sum (nu_model - nu*)^2 / sum nu*^2 of the trained network is below AMD's
and below 1 (no model), for the eddy viscosity and for the diffusivity."""
import numpy as np
import pytest

from diablo3_amd import train_closure

pytestmark = pytest.mark.gpu


def test_trained_closure_beats_amd_apriori():
    d, l = train_closure.front_pair((64, 33, 64), (16, 17, 16), re=5000.0, n2=1.0, amp=0.1, dt=1e-3)
    tr = []
    for seed in (0, 1):
        tr += train_closure.snapshots(d, l, seed, steps=200, every=50, log=lambda *a: None)
    te = train_closure.snapshots(d, l, 7, steps=200, every=50, log=lambda *a: None)
    X, T = train_closure.samples(tr)
    print(f"{X.shape[0]} samples; target means {T.mean(axis=0)}", flush=True)
    m, norm, hist = train_closure.train(X, T, epochs=40, batch=256, lr=2e-3, seed=3)
    print("epoch loss", hist[[0, 9, 19, 39]], flush=True)
    err = train_closure.apriori(m, te)
    untrained = train_closure.apriori(train_closure.closure.Mlp(seed=3), te)
    print("held-out relative MSE", {k: f"{v:.3f}" for k, v in err.items()},
          "untrained", {k: f"{v:.3f}" for k, v in untrained.items() if k.startswith("mlp")}, flush=True)
    assert hist[-1] < hist[0]
    assert err["mlp_nu"] < err["amd_nu"] and err["mlp_nu"] < 1.0
    assert err["mlp_kappa"] < err["amd_kappa"] and err["mlp_kappa"] < 1.0
