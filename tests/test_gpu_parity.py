"""Parity of the HIP step with the CPU oracle, through the C ABI.

Bars (written in each test): u1, u2, u3, theta and the pressure cp match the
oracle to relative L2 <= 1e-11 after 10 RKW3 steps (64x65x64 and 256x129x256);
transforms and the projection to ~1e-14; the closed-form RKW3 cases to 1e-14; the
energy-budget profiles of the GPU trajectory within 1e-10 of those of the oracle's
own trajectory; the mixed-precision path reported per field against the FP64
oracle and bounded at the error a float32 transform carries.

The pressure cp is recovered as phi / h_bar (for/solver.f90:1261-1267), so the
O(1e-16 |u|) round-off of the projected velocity's divergence reaches cp
amplified by 1/h_bar (7500 at dt = 1e-3, rk 2).  Each test prints the oracle's
disagreement with itself on cp under a legal round-off perturbation
(``noise_floor``: back substitution by reciprocal multiplication, or one transform
round trip of the initial state) beside the GPU's error; the bar stays fixed.
"""
import numpy as np
import pytest

import oracle
from diablo3_amd import configs
from diablo3_amd.config import FP64, MIXED
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def fields(s):
    return {"cu1": s.cu1, "cu2": s.cu2, "cu3": s.cu3, "cp": s.cp, **{f"cth{n}": t for n, t in enumerate(s.cth)}}


def compare(sg, so, tol, what="", floor=None, bars=None):
    """Relative L2 per field against a fixed bar (tol, or bars[field]); the oracle's
    own round-off disagreement (floor) is printed beside it, not used as a bar."""
    fg, fo = fields(sg), fields(so)
    errs = {k: rel(fg[k], fo[k]) for k in fo}
    bar = {k: (bars or {}).get(k, tol) for k in fo}
    print(what, {k: f"{v:.2e}" for k, v in errs.items()},
          "" if floor is None else {"oracle_self_floor": {k: f"{v:.1e}" for k, v in floor.items()}},
          flush=True)
    for k, e in errs.items():
        assert e <= bar[k], (what, k, e, bar[k])
    return errs


def noise_floor(p, s0, substeps, round_trip=True):
    """The oracle's disagreement with itself: per field, the larger relative change
    after `substeps` under (a) back substitution by reciprocal multiplication
    instead of division, (b) (round_trip) one fft_xz_to_physical +
    fft_xz_to_fourier of every field of the initial state.  Both are legal
    round-off perturbations of the algorithm."""
    dt = p.cfg.delta_t

    def run(state, variant):
        oracle.lib().oracle_set_variant(variant)
        try:
            o = oracle.Oracle(p, threads=16)
            o.upload_state(state)
            for n, rk in substeps:
                o.rk_substep(rk, time=(n - 1) * dt, time_step=n)
            return fields(o.download_state())
        finally:
            oracle.lib().oracle_set_variant(0)

    base, fb = run(s0, 0), run(s0, 1)
    out = {k: rel(fb[k], base[k]) for k in base}
    if round_trip:
        ot = oracle.Oracle(p, threads=16)
        s1 = s0.copy()
        for name in ("cu1", "cu2", "cu3"):
            setattr(s1, name, ot.fft_to_fourier(ot.fft_to_physical(getattr(s1, name))))
        s1.cth = [ot.fft_to_fourier(ot.fft_to_physical(t)) for t in s1.cth]
        fa = run(s1, 0)
        out = {k: max(out[k], rel(fa[k], base[k])) for k in base}
    return out


def divergence(p, s):
    """ikx u1 + (u2(j+1) - u2(j))/dyf(j) + ikz u3 on rows 1..Nyp."""
    cfg = p.cfg
    kx, kz, dyf, ny = p.kx[:cfg.kxr], p.kz, p.met["dyf"], cfg.ny
    return (1j * kx[None, None, :] * s.cu1[1:ny + 1]
            + (s.cu2[2:ny + 2] - s.cu2[1:ny + 1]) / dyf[1:ny + 1, None, None]
            + 1j * kz[None, :, None] * s.cu3[1:ny + 1])


# profiles of one kind (same units, the same term in another direction): a
# profile that vanishes identically in these flows (vme, the plane mean of v
# under the projection and the wall values; vv_dvdy; thv_m) holds only round-off,
# some 1e-28, and is measured against the largest profile of its kind
BUDGET_FAMILIES = [("ume", "vme", "wme"), ("urms", "vrms", "wrms"), ("uw", "uv", "wv"),
                   ("uu_dudx", "vv_dvdy", "wu_dwdx", "vu_dvdx", "uv_dudy", "wv_dwdy"),
                   ("dudy", "dwdy"), ("epsilon", "epsilon_m"), ("omega_x", "omega_y", "omega_z"),
                   ("thv", "thv_m")]


def budget_scale(so, name):
    base = name.rstrip("0123456789")
    suffix = name[len(base):]
    fam = next((f for f in BUDGET_FAMILIES if base in f), (base,))
    return max(np.abs(so[m + suffix]).max() for m in fam if m + suffix in so)


def compare_budget(g, o, what, bar=1e-10):
    """save_stats_chan's energy-budget profiles: the device statistics of the GPU
    state against the oracle's restatement on the oracle's own state, each
    profile's largest difference over the largest value of its kind."""
    from oracle import diagnostics
    sg = g.stats(time=0.0)
    so = diagnostics.save_stats(o, time=0.0)
    worst, wname = 0.0, ""
    for name in diagnostics.names(g.cfg.n_th):
        a, b = sg[name], so[name]
        scale = budget_scale(so, name)
        err = np.abs(a - b).max() / (scale if scale > 0 else 1.0)
        if err > worst:
            worst, wname = err, name
        assert err <= bar, (what, name, err)
    print(what, f"energy budget: worst profile {wname} rel err {worst:.2e} (bar {bar:.0e})", flush=True)
    return worst


@pytest.mark.parametrize("nx,ny,nz", [(16, 17, 16), (64, 33, 32), (32, 9, 128), (256, 5, 64), (512, 5, 16), (16, 5, 512), (1024, 5, 32), (32, 5, 1024)])
def test_transforms_match_oracle(nx, ny, nz):
    cfg = configs.small_dns(nx, ny, nz)
    p = Problem(cfg)
    g, o = Solver(p), oracle.Oracle(p)
    rng = np.random.default_rng(11)
    phys = rng.standard_normal((cfg.ny2, cfg.nz, cfg.nx))
    fg, fo = g.fft_to_fourier(phys), o.fft_to_fourier(phys)
    assert np.abs(fg - fo).max() < 1e-14 * np.abs(fo).max() * 10
    bg, bo = g.fft_to_physical(fo), o.fft_to_physical(fo)
    assert np.abs(bg - bo).max() < 1e-13 * np.abs(bo).max()


def test_rem_div_matches_oracle():
    cfg = configs.small_dns(32, 33, 32)
    p = Problem(cfg)
    g, o = Solver(p), oracle.Oracle(p)
    s0 = p.initial_state(seed=3, theta_noise=1e-3)
    g.upload_state(s0)
    o.upload_state(s0)
    g.rem_div()
    o.rem_div()
    compare(g.download_state(), o.download_state(), 1e-13, "rem_div")


@pytest.mark.parametrize("ny", [129, 257, 385])
def test_rem_div_tall_low_wavenumbers(ny):
    """The projection on the stretched grids of 256x129x256, 512x257x512 and
    1024x385x1024 (1/(dy dyf) up to 8e5 at the walls) with Lx = Lz = 8 pi, so that
    the lowest columns (kx^2 + kz^2 = 0.0625) are weakly dominant: the device's
    PCR against the oracle's Thomas sweep at 1e-11 on every field.  For scale, on
    this input: the oracle against itself with the rows eliminated top-down
    2e-13..9e-13; PCR in the row-sum form 6e-13..3e-12; plain PCR (the diagonal
    updated by b + al c + ga a) 3e-10, 1e-9, 8e-9 on cu2 at 129, 257, 385 rows."""
    import math
    cfg = configs.small_dns(16, ny, 32)
    cfg.lx = cfg.lz = 8.0 * math.pi
    p = Problem(cfg)
    g, o = Solver(p), oracle.Oracle(p)
    s0 = p.initial_state(seed=3, theta_noise=1e-3)
    rng = np.random.default_rng(7)
    for a in (s0.cu1, s0.cu2, s0.cu3):
        a[1:-1] += 1e-3 * (rng.standard_normal(a[1:-1].shape) + 1j * rng.standard_normal(a[1:-1].shape))
    g.upload_state(s0)
    o.upload_state(s0)
    g.rem_div()
    o.rem_div()
    compare(g.download_state(), o.download_state(), 1e-11, f"rem_div 16x{ny}x32, L 8 pi")


@pytest.mark.parametrize("nx,ny,nz", [(16, 17, 16), (32, 21, 64)])
def test_each_substep_matches_oracle(nx, ny, nz):
    cfg = configs.small_dns(nx, ny, nz)
    p = Problem(cfg)
    g, o = Solver(p), oracle.Oracle(p)
    s0 = p.initial_state(seed=4, theta_noise=1e-3)
    for a in (s0.cu1, s0.cu2, s0.cu3):
        a *= 200.0  # O(0.1) velocities so that the nonlinear terms are exercised
    o.upload_state(s0)
    o.rem_div()
    s0 = o.download_state()
    g.upload_state(s0)
    for rk in (1, 2, 3):
        g.rk_substep(rk, time_step=1)
        o.rk_substep(rk, time_step=1)
        fl = noise_floor(p, s0, [(1, r) for r in range(1, rk + 1)])
        compare(g.download_state(), o.download_state(), 1e-12, f"rk{rk}", fl, bars={"cp": 1e-11})
        cg, co = g.download_rk(), o.download_rk()
        # RK memory on the rows each equation owns
        js, je, ny_ = cfg.jstart, cfg.jend, cfg.ny
        assert rel(cg[0][0][js:je + 1], co[0][0][js:je + 1]) < 1e-12
        assert rel(cg[0][1][2:ny_ + 1], co[0][1][2:ny_ + 1]) < 1e-12
        assert rel(cg[0][2][js:je + 1], co[0][2][js:je + 1]) < 1e-12


def _oracle_run(cfg, steps, seed=0):
    po = Problem(cfg.copy(precision=FP64))
    o = oracle.Oracle(po, threads=16)
    s0 = po.initial_state(seed=seed, theta_noise=1e-3)
    o.upload_state(s0)
    o.rem_div()
    s0 = o.download_state()
    t = 0.0
    for n in range(steps):
        o.step(time=t, time_step=n + 1)
        t += cfg.delta_t
    return po, o, s0


def _gpu_run(cfg, s0, steps, precision=FP64):
    p = Problem(cfg.copy(precision=precision))
    g = Solver(p)
    g.upload_state(s0)
    t = 0.0
    for n in range(steps):
        g.step(time=t, time_step=n + 1)
        t += cfg.delta_t
    return p, g


def _run_pair(cfg, steps, seed=0, precision=FP64, floor=True):
    po, o, s0 = _oracle_run(cfg, steps, seed)
    p, g = _gpu_run(cfg, s0, steps, precision)
    fl = noise_floor(po, s0, [(n + 1, rk) for n in range(steps) for rk in (1, 2, 3)]) if floor else None
    return p, g, o, fl


def test_ten_steps_dns_64_fp64():
    """64x65x64 front DNS, 10 RKW3 steps: relative L2 <= 1e-11 on every field, cp
    included; divergence at round-off; energy-budget profiles within 1e-10 of the
    oracle's own trajectory's."""
    cfg = configs.dns_64()
    p, g, o, fl = _run_pair(cfg, 10)
    sg, so = g.download_state(), o.download_state()
    compare(sg, so, 1e-11, "dns_64 x10", fl)
    # discrete divergence at round-off on rows 2..Nyp-1
    d = divergence(p, sg)
    assert np.abs(d[1:-1]).max() < 1e-13 * np.abs(sg.cu1).max() * p.kx.max()
    compare_budget(g, o, "dns_64 x10")


def test_ten_steps_dns_256_fp64_and_mixed():
    """256x129x256 front DNS, 10 RKW3 steps from one seeded state: the FP64 path
    within 1e-11 of the oracle on every field, cp included, with the divergence at
    round-off and the energy-budget profiles within 1e-10; the mixed path (FP32
    transforms and products, FP64 solves, projection and mean mode) reported per
    field against the same FP64 oracle state and bounded at 1e-6."""
    cfg = configs.dns_256()
    po, o, s0 = _oracle_run(cfg, 10)
    so = o.download_state()
    fl = noise_floor(po, s0, [(n + 1, rk) for n in range(10) for rk in (1, 2, 3)], round_trip=False)
    p, g = _gpu_run(cfg, s0, 10, FP64)
    sg = g.download_state()
    compare(sg, so, 1e-11, "dns_256 fp64 x10", fl)
    d = divergence(p, sg)
    assert np.abs(d[1:-1]).max() < 1e-13 * np.abs(sg.cu1).max() * p.kx.max()
    compare_budget(g, o, "dns_256 fp64 x10")
    del g
    pm, gm = _gpu_run(cfg, s0, 10, MIXED)
    errs = compare(gm.download_state(), so, 1e-6, "dns_256 mixed x10 (error vs FP64 oracle)")
    assert max(errs.values()) > 1e-13  # it really ran in float32
    dm = divergence(pm, gm.download_state())
    assert np.abs(dm[1:-1]).max() < 1e-13 * np.abs(so.cu1).max() * p.kx.max()


def test_dns_256_configuration_length_mixed_vs_fp64():
    """256x129x256 front DNS over its configuration's 200 steps (dt 5e-4), FP64 and
    mixed from one projected state on the device: the mixed state within 1e-7 of the
    FP64 state on u and theta and 1e-6 on cp (relative L2; measured 2.4e-9 and
    3.2e-8), the mean buoyancy profile within 1e-12, and the divergence at
    round-off in both.  (200 oracle steps at this size would take minutes; the
    10-step test above holds the FP64 path to the oracle.)"""
    cfg = configs.dns_256()
    p = Problem(cfg)
    g0 = Solver(p)
    g0.upload_state(p.initial_state(seed=0, theta_noise=1e-3))
    g0.rem_div()
    s0 = g0.download_state()
    del g0
    p64, g64 = _gpu_run(cfg, s0, 200, FP64)
    s64 = g64.download_state()
    del g64
    pm, gm = _gpu_run(cfg, s0, 200, MIXED)
    sm = gm.download_state()
    del gm
    errs = {k: rel(a, b) for (k, a), b in zip(fields(sm).items(), fields(s64).values())}
    mt = rel(sm.cth[0][1:cfg.ny + 1, 0, 0].real, s64.cth[0][1:cfg.ny + 1, 0, 0].real)
    print("dns_256 x200 mixed vs FP64", {k: f"{v:.2e}" for k, v in errs.items()}, f"mean theta {mt:.2e}",
          flush=True)
    for k, e in errs.items():
        assert e <= (1e-6 if k == "cp" else 1e-7), (k, e)
    assert mt <= 1e-12
    for pp, ss in ((p64, s64), (pm, sm)):
        d = divergence(pp, ss)
        assert np.abs(d[1:-1]).max() < 1e-13 * np.abs(ss.cu1).max() * pp.kx.max()


@pytest.mark.parametrize("ny", [257, 385])
def test_tall_columns(ny):
    """128 x ny x 128 front DNS, 2 steps, FP64: the column kernels take the shapes
    the 512x257x512 and 1024x385x1024 grids select (k_project 4 columns per block
    with 5 and 7 rows per thread; k_cn 4 and 2 columns per block)."""
    cfg = configs.small_dns(128, ny, 128)
    p, g, o, _ = _run_pair(cfg, 2, seed=6, floor=False)
    compare(g.download_state(), o.download_state(), 1e-11, f"128x{ny}x128 x2")


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_wide_z_steps(precision):
    """16x17x512, 3 steps: the z transforms at nz = 512 take their widest column
    tiles (one 128-byte row segment per z row) in both precisions."""
    cfg = configs.small_dns(16, 17, 512)
    p, g, o, fl = _run_pair(cfg, 3, seed=5, precision=precision, floor=precision == FP64)
    if precision == FP64:
        compare(g.download_state(), o.download_state(), 1e-11, "wide z x3", fl)
    else:
        compare(g.download_state(), o.download_state(), 1e-4, "wide z mixed x3")


def test_mixed_precision_error_dns_64():
    cfg = configs.dns_64()
    p, g, o, _ = _run_pair(cfg, 10, precision=MIXED)
    errs = compare(g.download_state(), o.download_state(), 1e-4, "dns_64 mixed x10")
    assert max(errs.values()) > 1e-12  # it really ran in float32


# ---- closed forms on the device (independent of the oracle) ----
from test_oracle import _linear_cfg, _P, _zero_state  # noqa: E402


def test_gpu_rkw3_diffusion_polynomial():
    cfg = _linear_cfg(nx=16, nz=16, re=50.0, delta_t=0.05)
    p = Problem(cfg)
    g = Solver(p)
    s = _zero_state(cfg)
    s.cu1[:, 1, 0] = 0.3 + 0.1j
    s.cu1[:, cfg.twonkz, 0] = 0.3 - 0.1j
    g.upload_state(s)
    g.step(time_step=1)
    r = g.download_state()
    z = -cfg.nu * p.kz[1] ** 2 * cfg.delta_t
    assert np.abs(r.cu1[1:cfg.ny + 1, 1, 0] - (0.3 + 0.1j) * _P(z)).max() < 1e-14


def test_gpu_rkw3_coriolis_polynomial():
    cfg = _linear_cfg(nx=16, nz=16, ro=1.0 / 1.3, delta_t=0.1)
    p = Problem(cfg)
    g = Solver(p)
    s = _zero_state(cfg)
    s.cu1[:, 0, 0] = 0.7
    s.cu3[:, 0, 0] = -0.2
    g.upload_state(s)
    g.step(time_step=1)
    r = g.download_state()
    w = (0.7 - 0.2j) * _P(-1j * (cfg.ro_inv / cfg.delta) * cfg.delta_t)
    assert np.abs(r.cu1[1:cfg.ny + 1, 0, 0] - w.real).max() < 1e-14
    assert np.abs(r.cu3[1:cfg.ny + 1, 0, 0] - w.imag).max() < 1e-14
