"""A-posteriori evaluation of an LES closure against filtered DNS.

From a DNS of the front (this code, on the device), the state filtered onto the
LES grid at step n0 starts LES runs with each closure (les_model_type 5, AMD, and
6, the learned network); at later steps each LES state is compared with the
filtered DNS at the same time by
  * the mean buoyancy profile: theta_bar(y) = the (kx, kz) = (0, 0) mode of theta,
    relative L2 error over the interior rows;
  * the TKE spectrum: E(kx) = sum over y rows and kz of |u_i(kx, kz)|^2 with the
    mean mode removed, for kx = 1..Nkx, as the rms of log10(E_les / E_ref).

    python -m diablo3_amd.closure_eval --closure closure.npz --dns 256 129 256 --les 64 65 64
"""
from __future__ import annotations

import argparse
import json
import time as _time
from typing import Dict, List, Optional

import numpy as np

from . import closure, train_closure
from .problem import Problem
from .solver import Solver


def mean_profile(s, cfg) -> np.ndarray:
    return s.cth[0][1:cfg.ny, 0, 0].real.copy()


def tke_spectrum(s, cfg) -> np.ndarray:
    e = np.zeros(cfg.nkx + 1)
    for cu in (s.cu1, s.cu2, s.cu3):
        a = np.abs(cu[2:cfg.ny - 1]) ** 2
        a[:, 0, 0] = 0.0
        # kx > 0 appears twice in the real field (conjugate half)
        e += a.sum(axis=(0, 1))
    return e


def compare(s_les, s_ref, cfg) -> Dict[str, float]:
    pl, pr = mean_profile(s_les, cfg), mean_profile(s_ref, cfg)
    pr0 = pr - pr.mean()
    prof = float(np.linalg.norm(pl - pr) / max(np.linalg.norm(pr0), 1e-30))
    el, er = tke_spectrum(s_les, cfg)[1:], tke_spectrum(s_ref, cfg)[1:]
    ok = (er > 0) & (el > 0)
    spec = float(np.sqrt(np.mean(np.log10(el[ok] / er[ok]) ** 2))) if ok.any() else float("nan")
    return {"theta_profile_rel_l2": prof, "tke_spectrum_rms_log10": spec}


def filtered_states(dns_cfg, les_cfg, seed: int, steps: List[int], log=print):
    """The filtered DNS state on the LES grid at the given steps (ascending)."""
    pd, pl = Problem(dns_cfg), Problem(les_cfg)
    g = Solver(pd)
    g.upload_state(pd.initial_state(seed=seed, theta_noise=dns_cfg.kick / 2))
    g.rem_div()
    out = {}
    t = 0.0
    for n in range(1, max(steps) + 1):
        g.step(time=t, time_step=n, create_new_flow=True)
        t += dns_cfg.delta_t
        if n in steps:
            out[n] = train_closure.filtered(g, pd, pl).state
            log(f"  filtered DNS seed {seed} at step {n}")
    g.close()
    return out


def run_les(les_cfg, model: int, state, n0: int, steps: List[int], params=None, norm=None, sgs: bool = True):
    """LES from `state` at step n0; sgs False steps the same handle without the SGS
    model (the spin-up gate's DNS branch: create_new_flow with time_step 1)."""
    cfg = les_cfg.copy(les_model_type=model)
    p = Problem(cfg)
    g = Solver(p)
    if model == 6 and params is not None:
        g.set_closure(params, norm)
    g.upload_state(state)
    out = {}
    t = n0 * cfg.delta_t
    for n in range(n0 + 1, max(steps) + 1):
        if sgs:
            g.step(time=t, time_step=n, create_new_flow=False)
        else:
            g.step(time=t, time_step=1, create_new_flow=True)
        t += cfg.delta_t
        if n in steps:
            out[n] = g.download_state()
    g.close()
    return out


def evaluate(dns_cfg, les_cfg, seed: int, n0: int, steps: List[int], params, norm, log=print):
    ref = filtered_states(dns_cfg, les_cfg, seed, [n0] + steps, log)
    res = {"n0": n0, "steps": steps, "amd": {}, "learned": {}, "no_model": {}}
    runs = {"amd": run_les(les_cfg, 5, ref[n0], n0, steps),
            "learned": run_les(les_cfg, 6, ref[n0], n0, steps, params, norm),
            "no_model": run_les(les_cfg, 5, ref[n0], n0, steps, sgs=False)}
    for name, r in runs.items():
        for n in steps:
            res[name][n] = compare(r[n], ref[n], les_cfg)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--closure", default=None, help=".npz from train_closure (default: the committed one)")
    ap.add_argument("--dns", type=int, nargs=3, default=[256, 129, 256])
    ap.add_argument("--les", type=int, nargs=3, default=[64, 65, 64])
    ap.add_argument("--re", type=float, default=5000.0)
    ap.add_argument("--ri", type=float, default=1.0)
    ap.add_argument("--amp", type=float, default=0.1)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=99)
    ap.add_argument("--n0", type=int, default=400)
    ap.add_argument("--steps", type=int, nargs="+", default=[500, 600, 800])
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    params, norm, src = closure.load_params(a.closure)
    d, l = train_closure.front_pair(a.dns, a.les, a.re, a.ri, a.amp, a.dt)
    t0 = _time.time()
    res = evaluate(d, l, a.seed, a.n0, a.steps, params, norm)
    res["closure"] = src
    for n in a.steps:
        print(f"step {n}: AMD {res['amd'][n]}  learned {res['learned'][n]}  no model {res['no_model'][n]}",
              flush=True)
    print(f"({_time.time() - t0:.1f} s)")
    if a.out:
        json.dump(res, open(a.out, "w"), indent=1)


if __name__ == "__main__":
    main()
