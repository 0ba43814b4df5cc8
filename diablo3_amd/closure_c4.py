"""The learned-closure workload at its stated scale, in one process on the device.

Training data: seeded DNS of the finite front (front_les physics, Re 5000) at
512x257x512 for Ri 0.5, 1, 2, 4 with two seeds each; every `every` steps past
`start`, the state is box-filtered on the device (dgpu_box_filter: 4x in x and z,
the 129-row LES grid's cells in y) onto 128x129x128, and the features of
dgpu_les_features with the least-squares targets nu*, kappa* of the filtered
stresses become samples (about 2e8).  Training: the MLP 13-64-64-2 (GELU, BF16
MFMA with FP32 accumulation) for 20 epochs of Adam, lr 1e-3, on the device.
Evaluation on a held-out seed per Ri in `eval_ri`: a-priori relative errors of
the network and of AMD against nu*, kappa*; a-posteriori LES at 128x129x128
started from the filtered DNS at step n0 and run `les_steps` steps with the
learned closure (AMD fallback outside the training envelope), with AMD, and
with no model, each compared with the filtered DNS by the mean buoyancy profile
and the TKE spectrum (closure_eval.compare) every `compare_every` steps.

    python -m diablo3_amd.closure_c4 --out-dir c4_out
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time as _time

import numpy as np

from . import closure, closure_eval, sgs_data, train_closure
from .problem import Problem
from .solver import Solver


def log(msg, t0=[_time.time()]):
    print(f"[{_time.time() - t0[0]:7.1f} s] {msg}", flush=True)


def dns_run(d, l, seed: int, steps: int, start: int, every: int, keep=()):
    """One DNS from the seeded IC: (samples X, T per snapshot), a-priori
    snapshots, and the filtered states at the steps in `keep`."""
    pd, pl = Problem(d), Problem(l)
    g, gl = Solver(pd), Solver(pl)
    g.upload_state(pd.initial_state(seed=seed, theta_noise=d.kick / 2))
    g.rem_div()
    Xs, Ts, snaps, states = [], [], [], {}
    t = 0.0
    for n in range(1, steps + 1):
        g.step(time=t, time_step=n, create_new_flow=True)
        t += d.delta_t
        take = n > start and (n - start) % every == 0
        if take or n in keep:
            f = train_closure.filtered(g, pd, pl)
            if n in keep:
                states[n] = f.state
            if take:
                gl.upload_state(f.state)
                X, sc, am, _ = gl.les_features(t)
                snap = train_closure.Snapshot(X, sc, am, f.nu_star.ravel(), f.kappa_star.ravel(), seed, n)
                Xk, Tk, _ = sgs_data.make_samples(X, sc, None, snap.nu, snap.kappa)
                Xs.append(Xk)
                Ts.append(Tk)
                snaps.append(snap)
                log(f"  Ri {d.ri[0]} seed {seed} step {n} (t {t:.2f}): {Xk.shape[0]} samples, "
                    f"nu* mean {snap.nu.mean():.3e}, AMD nu_t mean {(sc * am[:, 0]).mean():.3e}")
    g.close()
    gl.close()
    return Xs, Ts, snaps, states


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--dns", type=int, nargs=3, default=[512, 257, 512])
    ap.add_argument("--les", type=int, nargs=3, default=[128, 129, 128])
    ap.add_argument("--re", type=float, default=5000.0)
    ap.add_argument("--ri", type=float, nargs="+", default=[0.5, 1.0, 2.0, 4.0])
    ap.add_argument("--seeds", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--length", type=float, default=2.0 * math.pi)
    ap.add_argument("--amp", type=float, default=0.1)
    ap.add_argument("--dt", type=float, default=4e-3)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--start", type=int, default=200)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--test-seed", type=int, default=99)
    ap.add_argument("--eval-ri", type=float, nargs="+", default=[0.5, 1.0])
    ap.add_argument("--n0", type=int, default=300)
    ap.add_argument("--les-steps", type=int, default=1000)
    ap.add_argument("--compare-every", type=int, default=250)
    ap.add_argument("--out-dir", default=".")
    a = ap.parse_args(argv)
    os.makedirs(a.out_dir, exist_ok=True)
    log(f"training data: DNS {a.dns} -> LES {a.les}, Ri {a.ri}, seeds {a.seeds}, L {a.length:.4f}, "
        f"steps {a.steps}, snapshots every {a.every} past {a.start}")
    Xs, Ts = [], []
    for ri in a.ri:
        d, l = train_closure.front_pair(a.dns, a.les, a.re, ri, a.amp, a.dt, a.length)
        for seed in a.seeds:
            X1, T1, _, _ = dns_run(d, l, seed, a.steps, a.start, a.every)
            Xs += X1
            Ts += T1
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    del Xs, Ts
    log(f"{X.shape[0]} training samples")
    m, norm, hist = train_closure.train(X, T, a.epochs, a.batch, a.lr)
    n_samples = int(X.shape[0])
    del X, T
    log(f"trained: epoch losses {np.array2string(hist, precision=4)}")
    res = {"dns": a.dns, "les": a.les, "re": a.re, "ri": a.ri, "seeds": a.seeds, "length": a.length,
           "amp": a.amp, "dt": a.dt, "steps": a.steps, "snapshots_every": a.every, "start": a.start,
           "samples": n_samples, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
           "epoch_loss": hist.tolist(), "eval": {}}
    checks = [a.n0 + k for k in range(a.compare_every, a.les_steps + 1, a.compare_every)]
    for ri in a.eval_ri:
        d, l = train_closure.front_pair(a.dns, a.les, a.re, ri, a.amp, a.dt, a.length)
        _, _, snaps, ref = dns_run(d, l, a.test_seed, max(checks), a.start, a.every, keep=[a.n0] + checks)
        err = train_closure.apriori(m, snaps)
        log(f"Ri {ri}: a-priori relative MSE {({k: round(v, 4) for k, v in err.items()})}")
        runs = {"learned": closure_eval.run_les(l, 6, ref[a.n0], a.n0, checks, m.params, norm),
                "amd": closure_eval.run_les(l, 5, ref[a.n0], a.n0, checks),
                "no_model": closure_eval.run_les(l, 5, ref[a.n0], a.n0, checks, sgs=False)}
        post = {name: {n: closure_eval.compare(r[n], ref[n], l) for n in checks} for name, r in runs.items()}
        for n in checks:
            log(f"Ri {ri} step {n}: " + "  ".join(f"{k} {v[n]['theta_profile_rel_l2']:.4f}/"
                                               f"{v[n]['tke_spectrum_rms_log10']:.4f}" for k, v in post.items()))
        res["eval"][str(ri)] = {"apriori": err, "aposteriori": {k: {str(n): v for n, v in p.items()}
                                                                 for k, p in post.items()},
                                "n0": a.n0, "checks": checks, "test_seed": a.test_seed}
    np.savez(os.path.join(a.out_dir, "closure_front.npz"), params=m.params, norm=norm, epoch_loss=hist)
    json.dump(res, open(os.path.join(a.out_dir, "closure_front_eval.json"), "w"), indent=1)
    log(f"wrote {a.out_dir}/closure_front.npz and closure_front_eval.json")


if __name__ == "__main__":
    main()
