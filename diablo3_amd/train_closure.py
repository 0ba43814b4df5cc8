"""The learned closure's workflow: DNS on the device -> box-filtered snapshots ->
device features and least-squares targets -> training on the device -> a
parameter block and normaliser for les_model_type 6.

    python -m diablo3_amd.train_closure --dns 128 65 128 --les 32 33 32 --seeds 0 1 \
        --steps 400 --snap 100 --epochs 20 --out closure.npz

The written .npz holds params (float32, DGPU_MLP_NPARAM), norm (DGPU_MLP_NNORM)
and the a-priori errors of the network and of AMD on held-out snapshots.
"""
from __future__ import annotations

import argparse
import math
import time as _time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import closure, configs, sgs_data
from .config import Config
from .problem import Problem
from .solver import Solver


def front_pair(dns_shape, les_shape, re: float, n2: float, amp: float, dt: float,
               length: float = 8.0 * math.pi):
    """A DNS config and its LES grid of the same finite front (front_les physics,
    Lx = Lz = length)."""
    d = configs.front_les(*dns_shape, re, dt, amp, n2=n2, length=length).copy(use_les=False)
    l = configs.front_les(*les_shape, re, dt, amp, n2=n2, length=length).copy(les_model_type=6)
    return d, l


def filtered(g: Solver, pd: Problem, pl: Problem) -> sgs_data.Filtered:
    """The DNS handle's state box-filtered onto the LES grid on the device
    (dgpu_box_filter), then targets and the filtered state on the host."""
    mf, mg = sgs_data.y_weights(pd.met, pl.met)
    f = g.box_filter(pd.cfg.nx // pl.cfg.nx, pd.cfg.nz // pl.cfg.nz, mf, mg)
    return sgs_data.targets(pl.cfg, pl.met, f)


@dataclass
class Snapshot:
    X: np.ndarray        # features [n][16]
    scale: np.ndarray    # C^2 D2 sqrt(2 s2) [n]
    amd: np.ndarray      # AMD velocity and scalar coefficients / sqrt(2 s2) [n][2]
    nu: np.ndarray       # nu* [n]
    kappa: np.ndarray    # kappa* [n]
    seed: int
    step: int


def snapshots(dns_cfg: Config, les_cfg: Config, seed: int, steps: int, every: int,
              start: int = 0, log=print) -> List[Snapshot]:
    """Run the DNS `steps` steps from a seeded IC; every `every` steps past
    `start`, filter, extract features on an LES handle, and form targets."""
    pd, pl = Problem(dns_cfg), Problem(les_cfg)
    g = Solver(pd)
    gl = Solver(pl)
    g.upload_state(pd.initial_state(seed=seed, theta_noise=dns_cfg.kick / 2))
    g.rem_div()
    out = []
    t = 0.0
    for n in range(1, steps + 1):
        g.step(time=t, time_step=n, create_new_flow=True)
        t += dns_cfg.delta_t
        if n > start and n % every == 0:
            f = filtered(g, pd, pl)
            gl.upload_state(f.state)
            X, sc, am, _ = gl.les_features(t)
            out.append(Snapshot(X, sc, am, f.nu_star.ravel(), f.kappa_star.ravel(), seed, n))
            log(f"  seed {seed} step {n}: {X.shape[0]} points, nu* mean {f.nu_star.mean():.3e}, "
                f"AMD nu_t mean {(sc * am[:, 0]).mean():.3e}")
    g.close()
    gl.close()
    return out


def samples(snaps: Sequence[Snapshot], cap_quantile: float = 0.995):
    Xs, Ts = [], []
    for s in snaps:
        X, T, _ = sgs_data.make_samples(s.X, s.scale, None, s.nu, s.kappa, cap_quantile)
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def apriori(m: closure.Mlp, snaps: Sequence[Snapshot]):
    """Relative mean-square errors of nu_t and kappa_t against nu*, kappa* on the
    given snapshots, for the network and for AMD (each over the points with a
    positive scale): sum (model - target)^2 / sum target^2."""
    num = {"mlp_nu": 0.0, "mlp_kappa": 0.0, "amd_nu": 0.0, "amd_kappa": 0.0}
    den = {"nu": 0.0, "kappa": 0.0}
    for s in snaps:
        ok = (s.scale > 0) & np.isfinite(s.nu) & np.isfinite(s.kappa)
        Y = m.predict(s.X[ok])
        sc = s.scale[ok].astype(np.float64)
        nu_m, ka_m = sc * np.maximum(0.0, Y[:, 0]), sc * np.maximum(0.0, Y[:, 1])
        nu_a, ka_a = sc * s.amd[ok, 0], sc * s.amd[ok, 1]
        nu_t, ka_t = s.nu[ok], s.kappa[ok]
        num["mlp_nu"] += np.sum((nu_m - nu_t) ** 2)
        num["mlp_kappa"] += np.sum((ka_m - ka_t) ** 2)
        num["amd_nu"] += np.sum((nu_a - nu_t) ** 2)
        num["amd_kappa"] += np.sum((ka_a - ka_t) ** 2)
        den["nu"] += np.sum(nu_t ** 2)
        den["kappa"] += np.sum(ka_t ** 2)
    return {k: v / den["nu" if k.endswith("nu") else "kappa"] for k, v in num.items()}


def train(X: np.ndarray, T: np.ndarray, epochs: int = 20, batch: int = 1024, lr: float = 1e-3,
          seed: int = 0, envelope_quantile: float = 0.0):
    """Normaliser from X, loss weights 1/var(T_o), Adam on the device."""
    norm = closure.make_norm(X, envelope_quantile)
    var = T.var(axis=0)
    lw = (1.0 / max(var[0], 1e-30), 1.0 / max(var[1], 1e-30))
    m = closure.Mlp(seed=seed)
    m.set_norm(norm)
    m.set_data(X, T)
    hist = m.train(epochs, batch, lr, seed, lw)
    return m, norm, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--dns", type=int, nargs=3, default=[128, 65, 128])
    ap.add_argument("--les", type=int, nargs=3, default=[32, 33, 32])
    ap.add_argument("--re", type=float, default=5000.0)
    ap.add_argument("--ri", type=float, nargs="+", default=[1.0])
    ap.add_argument("--amp", type=float, default=0.1)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--seeds", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--test-seed", type=int, default=99)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--snap", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="closure.npz")
    a = ap.parse_args(argv)
    t0 = _time.time()
    train_snaps, test_snaps = [], []
    for ri in a.ri:
        d, l = front_pair(a.dns, a.les, a.re, ri, a.amp, a.dt)
        for s in a.seeds:
            train_snaps += snapshots(d, l, s, a.steps, a.snap)
        test_snaps += snapshots(d, l, a.test_seed, a.steps, a.snap)
    X, T = samples(train_snaps)
    print(f"{X.shape[0]} training samples ({_time.time() - t0:.1f} s)", flush=True)
    m, norm, hist = train(X, T, a.epochs, a.batch, a.lr)
    print("epoch losses", np.array2string(hist, precision=4), flush=True)
    err = apriori(m, test_snaps)
    print("a-priori relative MSE on held-out snapshots:", {k: round(v, 4) for k, v in err.items()}, flush=True)
    np.savez(a.out, params=m.params, norm=norm, epoch_loss=hist, **{f"err_{k}": v for k, v in err.items()})
    print(f"wrote {a.out} ({_time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
