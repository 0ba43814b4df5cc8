#!/usr/bin/env python3
"""RKW3 + Crank–Nicolson step throughput on one MI355X.

Workload: the plain-DNS infinite front at 256x129x256 (code x, y, z), FP64 by
default, synthetic seeded initial condition, resident in HBM before timing starts.
One "step" is the full RKW3 step (three rk_chan_1 substeps) through the C ABI.

Printed JSON line: value = Nx*Ny*Nz / t_step in Mpoint-updates/s; roofline for the
dominant kernel (algorithmic bytes per launch / its mean HIP-event duration over
the timed region); cpu_baseline = the CPU oracle timed on this host's cores on a
bounded sample (the median substep of three steps on all cores, and on one
core); learned_vs_amd_les = the learned-closure mixed LES step against the
FP64 and mixed AMD LES steps at 512x257x512 (the metric's third part).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RKW3 step Mpoint-updates/s, %HBM peak; learned-vs-AMD LES speedup at fixed error"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(cfg, mixed):
    """Algorithmic HBM bytes per launch of each kernel (each array read or written once).

    C = one packed Fourier field (complex double), H = one x-spectrum field
    (complex T); nf = 3 + n_th state fields, no = 6 + 2 n_th product spectra,
    nj = planes that carry products."""
    nf, no = 3 + cfg.n_th, 6 + 2 * cfg.n_th
    s = 4 if mixed else 8
    ny2 = cfg.ny2
    jlo = min([cfg.jstart, 2] + list(cfg.jstart_th[:cfg.n_th]))
    nj = cfg.ny - jlo + 1
    C = cfg.kxr * cfg.kzr * 16 * ny2
    Cj = cfg.kxr * cfg.kzr * 16 * nj
    H = cfg.kxr * cfg.nz * 2 * s * ny2
    Hj = cfg.kxr * cfg.nz * 2 * s * nj
    Cj_t = cfg.kxr * cfg.kzr * 2 * s * nj   # nonlinear terms, complex T
    Rt = cfg.nx * cfg.nz * s * ny2
    Rtj = cfg.nx * cfg.nz * s * nj
    out = {
        "k_inv_z": nf * C + nf * H,
        # DNS x pass marching in y: the x-spectra of planes jlo-1..ny+1 in, the
        # product spectra of planes jlo..ny out (physical fields stay on chip)
        "k_xmarch": nf * cfg.kxr * cfg.nz * 2 * s * (nj + 2) + no * Hj,
        # LES x pass: physical fields (and SGS fields) in, product spectra out
        "k_xprod": nf * Rtj + no * Hj,
        "k_fwdz_nl": no * Hj + nf * Cj_t,
        # nonlinear terms, state, cp and (rk 2, 3: 2 of 3 launches) old cf in; cf and state out
        "k_cn": nf * Cj_t + (nf + 1) * C + (2.0 / 3.0) * nf * Cj + nf * Cj + nf * C,
        "k_project": 4 * C + 4 * C,
    }
    if cfg.use_les:
        # physical fields: R8 FP64, Rt in the transform precision; na SGS addition fields
        R8 = cfg.nx * cfg.nz * 8 * ny2
        Hf = cfg.kxr * cfg.nz * 2 * 4 * ny2        # float x-spectrum (increment transforms)
        Rf = cfg.nx * cfg.nz * 4 * ny2
        na = 7 + 2 * cfg.n_th
        learned = cfg.les_model_type == 6
        out["k_bc_post"] = nf * 6 * C / ny2
        # z inverse of v and dv/dz (2 nf), x inverse of v, dv/dx, dv/dz (3 nf)
        out["les_grad"] = 2 * nf * C + 2 * nf * H + 3 * nf * H + 3 * nf * Rt
        # gradient fields in; nu_t, kappa_t, SGS fields out (the learned model also
        # reads its four coefficient planes)
        out["k_amd"] = 3 * nf * Rt + (1 + cfg.n_th + na) * Rt + (4 * Rt if learned else 0)
        if learned:
            out["k_closure"] = 3 * nf * Rt + 4 * Rt
        out["k_xprod"] += na * Rtj
        # assembly only: the right-hand side goes out instead of the solve
        out["k_cn"] = nf * Cj_t + (nf + 1) * C + (2.0 / 3.0) * nf * Cj + nf * Cj + nf * C
        npts = cfg.nx * cfg.nz * ny2
        if mixed:
            # increments: u (float), r - u (float), nu in (twice); 1/b', g' (float) out and
            # back; delta out
            out["k_cn_phys"] = nf * npts * (4 + 4 + 2 * s + 8 + 8 + 4)
            # per launch (two per substep): C -> Hf -> Rf, then Rf -> Hf -> C (read + write)
            out["les_cn_fft"] = 0.5 * nf * ((C + 2 * Hf + Rf) + (Rf + 2 * Hf + 2 * C))
        else:
            out["k_cn_phys"] = nf * npts * (8 + 8 + 2 * s + 16 + 16 + 8)
            out["les_cn_fft"] = 0.5 * nf * ((C + 2 * H + R8) + (R8 + 2 * H + C))
    return out


def model_v1_bytes_per_step(cfg, s):
    """The step-level 'canonical algorithm model v1': 30R + 30H + 57C per DNS
    substep, 96R + 88H + 89C per LES substep."""
    yg = cfg.ny + 2
    R = cfg.nx * cfg.nz * yg * s
    H = cfg.kxr * cfg.nz * yg * 2 * s
    C = cfg.kxr * cfg.kzr * yg * 2 * s
    if cfg.use_les:
        return 3 * (96 * R + 88 * H + 89 * C)
    return 3 * (30 * R + 30 * H + 57 * C)


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/*), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


TIMING_EVERY = 10


def log(msg):
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(cfg, threads, steps):
    """The oracle on this host's cores, timed per substep: the median substep
    time x 3 over `steps` RKW3 steps (every rk_step costs the same: the same
    transforms and solves), the reference's own per-iteration wall time
    (for/diablo.f90:104-109) taken as a median."""
    import statistics
    import oracle
    from diablo3_amd.problem import Problem
    p = Problem(cfg.copy(precision=0))
    o = oracle.Oracle(p, threads=threads)
    o.upload_state(p.initial_state(seed=0, theta_noise=1e-3))
    o.rem_div()
    cnf = 0 if cfg.use_les else 1
    times = []
    for n in range(steps):
        for rk in (1, 2, 3):
            t0 = time.perf_counter()
            o.rk_substep(rk, time=n * cfg.delta_t, time_step=n + 1, create_new_flow=cnf)
            times.append(time.perf_counter() - t0)
    dt = 3.0 * statistics.median(times)
    return {"value": round(cfg.nx * cfg.ny * cfg.nz / dt / 1e6, 3), "unit": "Mpoint-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {len(times)} substeps x 3 ({steps} RKW3 step(s)) of the same "
                      f"{cfg.nx}x{cfg.ny}x{cfg.nz} FP64 workload, C restatement of rk_chan_1 "
                      f"(physical-space CN, {59 if cfg.use_les else 23} transforms/substep), "
                      f"{'OpenMP ' if threads > 1 else ''}{threads} thread(s), {dt:.2f} s/step"}


def copy_calibration(nbytes, reps=3):
    """A plain device copy of known size (read nbytes, write nbytes per launch):
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters
    to bytes."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": reps,
                      "GB/s_read_plus_write": round(2 * nbytes / t / 1e9, 1)}), flush=True)


def les_speedup(steps, warmup):
    """The metric's third part on the 512x257x512 finite-front LES: ms per step of
    the learned-closure step in mixed precision against the plain FP64 AMD step
    (and the mixed AMD step), each from the same seeded state, SGS on from step 1."""
    from diablo3_amd import closure, configs
    from diablo3_amd.config import FP64, MIXED
    from diablo3_amd.problem import Problem
    from diablo3_amd.solver import Solver
    base = configs.les_512(FP64)
    p0 = Problem(base)
    s0 = p0.initial_state(seed=0, theta_noise=1e-3)
    params, norm, src = closure.load_params()
    out = {}
    for name, prec, model in (("fp64_amd", FP64, 5), ("mixed_amd", MIXED, 5), ("mixed_learned", MIXED, 6)):
        p = Problem(base.copy(precision=prec, les_model_type=model))
        g = Solver(p)
        if model == 6:
            g.set_closure(params, norm)
        g.upload_state(s0)
        g.rem_div()
        t = 0.0
        for n in range(warmup):
            g.step(time=t, time_step=n + 1, create_new_flow=False)
            t += base.delta_t
        g.synchronize()
        t0 = time.perf_counter()
        for n in range(steps):
            g.step(time=t, time_step=warmup + n + 1, create_new_flow=False)
            t += base.delta_t
        g.synchronize()
        out[name + "_ms_per_step"] = round((time.perf_counter() - t0) / steps * 1e3, 3)
        if model == 6:
            # the share of the closure's points that took the network in the last
            # substep (the rest left the training envelope and took AMD)
            nn, npts = g.closure_counts()
            out["learned_network_fraction"] = round(nn / max(npts, 1), 4)
        g.close()
        log(f"LES 512 {name}: {out[name + '_ms_per_step']} ms/step")
    out["speedup_learned_mixed_vs_fp64_amd"] = round(out["fp64_amd_ms_per_step"] / out["mixed_learned_ms_per_step"], 3)
    out["speedup_mixed_amd_vs_fp64_amd"] = round(out["fp64_amd_ms_per_step"] / out["mixed_amd_ms_per_step"], 3)
    out["workload"] = (f"front LES {base.nx}x{base.ny}x{base.nz}, Re {base.re}, dt {base.delta_t}, "
                       f"{steps} timed steps after {warmup}; learned closure {src}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dns_256", choices=["dns_256", "dns_64", "les_512", "les_128", "les_1024"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--closure", default="amd", choices=["amd", "learned"],
                    help="LES workloads: the AMD closure (type 5) or the learned one (type 6)")
    ap.add_argument("--closure-file", default=None,
                    help="learned closure parameters (.npz with params, norm); default: the "
                         "committed front closure if present, else a seeded network")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-les", action="store_true",
                    help="skip the learned-vs-AMD LES leg (512x257x512) of the default DNS run")
    ap.add_argument("--les-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (rocprofv3 counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only (no collectives, no sharding); "
              "use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from diablo3_amd import configs
    from diablo3_amd.config import FP64, MIXED
    from diablo3_amd.problem import Problem
    from diablo3_amd.solver import Solver

    mixed = args.precision == "mixed"
    cfg = configs.NAMED[args.workload](MIXED if mixed else FP64)
    learned = args.closure == "learned"
    if learned:
        if not cfg.use_les:
            print("bench.py: --closure learned needs an LES workload", file=sys.stderr)
            sys.exit(2)
        cfg = cfg.copy(les_model_type=6)
    p = Problem(cfg)
    g = Solver(p)
    closure_src = None
    if learned:
        from diablo3_amd import closure
        params, norm, closure_src = closure.load_params(args.closure_file)
        g.set_closure(params, norm)
    g.upload_state(p.initial_state(seed=0, theta_noise=1e-3))
    g.rem_div()
    # LES: a restart-style run (create_new_flow false), so the SGS model is on from step 1
    cnf = not cfg.use_les
    t = 0.0
    for n in range(args.warmup):
        g.step(time=t, time_step=n + 1, create_new_flow=cnf)
        t += cfg.delta_t
    g.timing(True)
    # kernel events on one step in TIMING_EVERY: the per-kernel durations come from
    # those steps, and the timed loop carries the events' cost (36 records per
    # step, about 6% of a 256x129x256 step) on them only
    g.timing_every(TIMING_EVERY)
    g.synchronize()
    t0 = time.perf_counter()
    for n in range(args.steps):
        g.step(time=t, time_step=args.warmup + n + 1, create_new_flow=cnf)
        t += cfg.delta_t
    g.synchronize()
    t1 = time.perf_counter()
    ktimes = g.timing_read()
    g.timing(False)
    ms = (t1 - t0) / args.steps * 1e3
    points = cfg.nx * cfg.ny * cfg.nz
    value = points / (ms * 1e-3) / 1e6

    kb = kernel_bytes(cfg, mixed)
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    tot, n = ktimes[dom]
    avg_s = tot / n * 1e-3
    achieved = kb[dom] / avg_s / 1e9
    workload = f"{args.workload}_{args.precision}"
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom, workload),
            "kernel": dom, "algorithmic_bytes_per_launch": int(kb[dom]),
            "mean_launch_ms": round(avg_s * 1e3, 5)}
    per_kernel = {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 5), "launches": v[1],
                      "GB/s": round(kb[k] / (v[0] / max(v[1], 1) * 1e-3) / 1e9, 1)}
                  for k, v in ktimes.items()}
    s = 4 if mixed else 8
    step_model = model_v1_bytes_per_step(cfg, s)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpoint-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if not mixed else "f32+f64",
        "data": "synthetic (seeded front IC, projected divergence-free)",
        "config": {"workload": f"front {('LES (learned closure)' if learned else 'LES (AMD closure)') if cfg.use_les else 'DNS'} "
                               f"{cfg.nx}x{cfg.ny}x{cfg.nz} (x,y,z), "
                               f"{'FP64' if not mixed else 'mixed FP32 transforms / FP64 solves'}, "
                               f"dt {cfg.delta_t}, Re {cfg.re}",
                   "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz, "precision": args.precision,
                   **({"closure": closure_src} if learned else {})},
        "roofline": roof,
        "step_hbm_frac_model_v1": round(step_model / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": per_kernel,
    }
    g.close()
    log(f"{args.workload} {args.precision}: {ms:.4f} ms/step over {args.steps} steps")
    if not args.no_les and args.workload == "dns_256":
        out["learned_vs_amd_les"] = les_speedup(args.les_steps, 2)
        log(f"LES leg: {out['learned_vs_amd_les']}")
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cfg64 = configs.NAMED[args.workload](FP64)
        out["cpu_baseline"] = cpu_baseline(cfg64, threads, steps=3)
        log(f"CPU baseline, {threads} threads: {out['cpu_baseline']['sample']}")
        # one core, as the reference's one-rank run: the median over three steps
        out["cpu_baseline"]["one_core"] = cpu_baseline(cfg64, 1, steps=3)
        log(f"CPU baseline, 1 thread: {out['cpu_baseline']['one_core']['sample']}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
