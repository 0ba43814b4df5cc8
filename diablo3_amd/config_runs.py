"""The DNS and LES workloads at the lengths their configurations state, on the
device: FP64 against the mixed-precision split from the same projected initial
state, and at 1024x385x1024 the learned closure against AMD.

* front DNS 256x129x256, 200 steps at dt 5e-4, FP64 and mixed;
* finite-front LES 512x257x512 (AMD), 500 steps at dt 2e-4, FP64 and mixed;
* finite-front LES 1024x385x1024, 200 steps at dt 1e-4, mixed, AMD and the
  committed learned closure (AMD outside its envelope).

Each run reports ms per step (wall time of the step loop over the device), the
final discrete divergence (rows 2..Nyp-1, over max|u1| kx_max), and per pair of
runs the relative L2 difference of every Fourier field and of the mean buoyancy
profile (the kx = kz = 0 mode of th, rows 1..Nyp) at the end.  The SGS model is
on from step 1 (create_new_flow false), as in the benchmark.

    python -m diablo3_amd.config_runs --out config_runs.json
"""
from __future__ import annotations

import argparse
import json
import time as _time

import numpy as np

from . import closure, configs
from .config import FP64, MIXED
from .problem import Problem
from .solver import Solver


def log(msg, t0=[_time.time()]):
    print(f"[{_time.time() - t0[0]:7.1f} s] {msg}", flush=True)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def fields(s):
    return {"cu1": s.cu1, "cu2": s.cu2, "cu3": s.cu3, "cp": s.cp, **{f"cth{n}": t for n, t in enumerate(s.cth)}}


def divergence(p, s):
    """max |div u| on rows 2..Nyp-1 over max|u1| kx_max (for/flow.f90:487 checks)."""
    cfg = p.cfg
    kx, kz, dyf, ny = p.kx[:cfg.kxr], p.kz, p.met["dyf"], cfg.ny
    d = (1j * kx[None, None, :] * s.cu1[1:ny + 1]
         + (s.cu2[2:ny + 2] - s.cu2[1:ny + 1]) / dyf[1:ny + 1, None, None]
         + 1j * kz[None, :, None] * s.cu3[1:ny + 1])
    return float(np.abs(d[1:-1]).max() / (np.abs(s.cu1).max() * kx.max()))


def mean_theta(p, s):
    return s.cth[0][1:p.cfg.ny + 1, 0, 0].real.copy()


def projected_state(cfg, seed=0):
    p = Problem(cfg)
    s = p.initial_state(seed=seed, theta_noise=1e-3)
    g = Solver(p)
    g.upload_state(s)
    g.rem_div()
    s0 = g.download_state()
    g.close()
    return s0


def run(cfg, s0, steps, label, closure_params=None):
    p = Problem(cfg)
    g = Solver(p)
    if closure_params is not None:
        g.set_closure(*closure_params)
    g.upload_state(s0)
    t = 0.0
    g.synchronize()
    t0 = _time.perf_counter()
    for n in range(1, steps + 1):
        g.step(time=t, time_step=n, create_new_flow=False)
        t += cfg.delta_t
        if n % 50 == 0:
            g.synchronize()
            log(f"{label}: step {n}")
    g.synchronize()
    ms = (_time.perf_counter() - t0) / steps * 1e3
    out = {"ms_per_step": round(ms, 3), "steps": steps}
    if closure_params is not None:
        nn, npts = g.closure_counts()
        out["network_fraction"] = round(nn / max(npts, 1), 4)
    s = g.download_state()
    out["divergence"] = divergence(p, s)
    out["finite"] = bool(all(np.all(np.isfinite(a)) for a in fields(s).values()))
    if cfg.use_les:
        nut, kt = g.download_les()
        out["nu_t_max"] = float(nut.max())
        out["nu_t_min"] = float(nut.min())
    g.close()
    log(f"{label}: {out}")
    return out, p, s


def compare(pa, sa, pb, sb):
    fa, fb = fields(sa), fields(sb)
    out = {k: rel(fa[k], fb[k]) for k in fb}
    out["mean_theta"] = rel(mean_theta(pa, sa), mean_theta(pb, sb))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="config_runs.json")
    ap.add_argument("--dns-steps", type=int, default=200)
    ap.add_argument("--les-steps", type=int, default=500)
    ap.add_argument("--big-steps", type=int, default=200)
    args = ap.parse_args(argv)
    res = {}

    log("DNS 256x129x256: initial state")
    c64 = configs.dns_256(FP64)
    s0 = projected_state(c64)
    r64, p64, s64 = run(c64, s0, args.dns_steps, "DNS 256 FP64")
    rmx, pmx, smx = run(configs.dns_256(MIXED), s0, args.dns_steps, "DNS 256 mixed")
    res["dns_256"] = {"fp64": r64, "mixed": rmx, "mixed_vs_fp64": compare(pmx, smx, p64, s64)}
    log(f"DNS 256 mixed vs FP64: {res['dns_256']['mixed_vs_fp64']}")
    del s64, smx

    log("LES 512x257x512: initial state")
    c64 = configs.les_512(FP64)
    s0 = projected_state(c64)
    r64, p64, s64 = run(c64, s0, args.les_steps, "LES 512 FP64")
    rmx, pmx, smx = run(configs.les_512(MIXED), s0, args.les_steps, "LES 512 mixed")
    res["les_512"] = {"fp64": r64, "mixed": rmx, "mixed_vs_fp64": compare(pmx, smx, p64, s64)}
    log(f"LES 512 mixed vs FP64: {res['les_512']['mixed_vs_fp64']}")
    del s64, smx, s0

    log("LES 1024x385x1024: initial state")
    cmx = configs.les_1024(MIXED)
    s0 = projected_state(cmx)
    params, norm, src = closure.load_params()
    ra, pa, sa = run(cmx, s0, args.big_steps, "LES 1024 mixed AMD")
    rl, pl, sl = run(cmx.copy(les_model_type=6), s0, args.big_steps, "LES 1024 mixed learned", (params, norm))
    res["les_1024"] = {"mixed_amd": ra, "mixed_learned": rl, "closure": src,
                       "learned_vs_amd": compare(pl, sl, pa, sa),
                       "speedup_learned_vs_amd": round(ra["ms_per_step"] / rl["ms_per_step"], 3)}
    log(f"LES 1024 learned vs AMD: {res['les_1024']['learned_vs_amd']}")
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
