"""Device energy-budget statistics (dgpu_stats) against the oracle's restatement
of save_stats_chan, through the C ABI.

Bar: every profile within 1e-10 of the oracle, relative to the profile's largest
magnitude (the acceptance bar for the budget terms); the state after the call
equals the oracle's state after the same BC posts (save_stats_chan applies them to
the flow it saves).  Both averaging modes: plane means (x-homogeneous infinite
front) and z-line means (finite front, IC type 6).
"""
import numpy as np
import pytest

import oracle
from oracle import diagnostics
from diablo3_amd import configs
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver

pytestmark = pytest.mark.gpu


def prepared(cfg, seed=3):
    p = Problem(cfg)
    o = oracle.Oracle(p, threads=16)
    s = p.initial_state(seed=seed, theta_noise=1e-2)
    rng = np.random.default_rng(seed)
    sh = (cfg.ny - 3, 4, 4)
    for a in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        a[2:cfg.ny - 1, 1:5, 1:5] += 0.02 * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    o.upload_state(s)
    o.rem_div()
    for rk in (1, 2, 3):
        o.rk_substep(rk, time=0.0, time_step=1)
    return p, o


@pytest.mark.parametrize("which", ["infinite_front_dns", "finite_front"])
def test_stats_match_oracle(which):
    if which == "infinite_front_dns":
        cfg = configs.small_dns(32, 33, 32)
    else:
        cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
        cfg.use_les = False  # the state is advanced by the DNS branch here
    p, o = prepared(cfg)
    s0 = o.download_state()
    g = Solver(p)
    g.upload_state(s0)
    sg = g.stats(time=0.0)
    so = diagnostics.save_stats(o, time=0.0)
    print(which, "homogeneous_x", diagnostics.homogeneous_x(cfg), flush=True)
    worst = 0.0
    for name in diagnostics.names(cfg.n_th):
        a, b = sg[name], so[name]
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / (scale if scale > 0 else 1.0)
        worst = max(worst, err)
        print(f"  {name:10s} max|ref| {scale:.3e} rel err {err:.2e}", flush=True)
        assert err <= 1e-10, (name, err)
    assert abs(sg["u1z_x0"] - so["u1z_x0"]) <= 1e-10 * max(abs(so["u1z_x0"]), 1e-300)
    ftx_err = np.abs(sg["FTx_uu"] - so["FTx_uu"]).max() / np.abs(so["FTx_uu"]).max()
    print(f"  FTx_uu     max|ref| {np.abs(so['FTx_uu']).max():.3e} rel err {ftx_err:.2e}", flush=True)
    assert sg["FTx_uu"].shape == (cfg.nx // 2,) and ftx_err <= 1e-10
    # the side effect on the state: the BC posts
    a, b = g.download_state(), o.download_state()
    for x, y in ((a.cu1, b.cu1), (a.cu2, b.cu2), (a.cu3, b.cu3), (a.cth[0], b.cth[0])):
        assert np.array_equal(x, y)
    # the budget terms are not trivially zero on this state
    assert so["epsilon"].max() > 0 and np.abs(so["thv01"]).max() > 0 and so["urms"].max() > 0
    if not diagnostics.homogeneous_x(cfg):
        assert so["epsilon_m"].max() > 0 and np.abs(so["uu_dudx"]).max() > 0


@pytest.mark.parametrize("which", ["infinite_front_dns", "finite_front"])
def test_zstar_matches_oracle(which):
    """dgpu_stats_zstar (compute_BPE, Bin_Ystar_and_Write) against the oracle:
    PDF, BPE and every binned mean within 1e-10 (relative to the largest
    magnitude; NaN exactly where a bin is empty), the bin edges to 1e-15."""
    if which == "infinite_front_dns":
        cfg = configs.small_dns(32, 33, 32)
    else:
        cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
        cfg.use_les = False
    p, o = prepared(cfg)
    g = Solver(p)
    g.upload_state(o.download_state())
    zg = g.stats_zstar(time=0.0)
    zo = diagnostics.zstar(o, time=0.0)
    occupied = int(np.count_nonzero(zo["th1PDF"]))
    print(which, f"BPE {zg['BPE']:.6e} / {zo['BPE']:.6e}, {occupied} of {cfg.nx} bins occupied", flush=True)
    assert occupied > 2
    assert np.abs(zg["th1bin"] - zo["th1bin"]).max() <= 1e-15 * np.abs(zo["th1bin"]).max()
    assert abs(zg["BPE"] - zo["BPE"]) <= 1e-10 * abs(zo["BPE"])
    for name in ["th1PDF"] + diagnostics.ZSTAR_FIELDS + ["thw_zstar01", "chi_zstar01"]:
        a, b = zg[name], zo[name]
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        m = ~np.isnan(b)
        scale = np.abs(b[m]).max()
        err = np.abs(a[m] - b[m]).max() / (scale if scale > 0 else 1.0)
        print(f"  {name:14s} max|ref| {scale:.3e} rel err {err:.2e}", flush=True)
        assert err <= 1e-10, (name, err)


@pytest.mark.parametrize("which", ["infinite_front_dns", "finite_front", "finite_front_les"])
def test_movie_matches_oracle(which):
    """dgpu_stats_movie: the mean_xz.h5 planes save_stats_chan writes for the
    configuration and every movie.h5 slice, within 1e-10 of the oracle (relative
    to each array's largest magnitude, with a 1e-8 floor for means at round-off).
    The LES case compares after one LES step of both from the same state (nu_t)."""
    from diablo3_amd import stats_io
    if which == "infinite_front_dns":
        cfg = configs.small_dns(32, 33, 32)
    else:
        cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
        cfg.use_les = which == "finite_front_les"
    p, o = prepared(cfg) if not cfg.use_les else (Problem(cfg), None)
    if cfg.use_les:
        o = oracle.Oracle(p, threads=16)
        o.upload_state(p.initial_state(seed=4, theta_noise=1e-2))
        o.rem_div()
    g = Solver(p)
    g.upload_state(o.download_state())
    if cfg.use_les:  # past the LES spin-up gate, so nu_t is live
        g.step(time=0.0, time_step=1, create_new_flow=False)
        o.step(time=0.0, time_step=1, create_new_flow=0)
    nxm, nym, nzm = stats_io.movie_indices(cfg, p.met["gyf"], 0.3 * cfg.lx, 0.25, 0.6 * cfg.lz)
    print(which, "movie indices", nxm, nym, nzm, flush=True)
    pg, sg = g.stats_movie(0.0, nxm, nym, nzm)
    po, so = diagnostics.movie(o, nxm, nym, nzm)
    for name in stats_io.written_planes(cfg):
        a, b = pg[name], po[name]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-8)
        print(f"  {name:13s} max|ref| {np.abs(b).max():.3e} rel err {err:.2e}", flush=True)
        assert a.shape == (cfg.ny, cfg.nx) and err <= 1e-10, (name, err)
    assert set(sg) == set(so)
    for name in sorted(so):
        a, b = sg[name], so[name]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-8)
        print(f"  {name:13s} max|ref| {np.abs(b).max():.3e} rel err {err:.2e}", flush=True)
        assert a.shape == b.shape and err <= 1e-10, (name, err)
    assert (np.abs(so["nu_t_xz"]).max() > 0) == cfg.use_les
    g.close()


def test_movie_index_checks_and_missing_row():
    """dgpu_stats_movie refuses slice indices outside the grid (before any launch)
    and, for NyMovie = -1 (the row not found), returns no x-z slices."""
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    g = Solver(p)
    g.upload_state(p.initial_state(seed=1, theta_noise=1e-2))
    g.rem_div()
    for bad in ((cfg.nx, 3, 0), (0, 0, 0), (0, cfg.ny + 1, 0), (0, 3, cfg.nz), (-1, 3, 0)):
        with pytest.raises(RuntimeError, match="movie indices"):
            g.stats_movie(0.0, *bad)
    planes, sl = g.stats_movie(0.0, 1, -1, 2)
    assert "u_xz" in sl and "th1_yz" in sl and not any(k.endswith("_xy") for k in sl)
    g.close()


def test_movie_nu_t_slices_mixed_precision():
    """The nu_t slices of a mixed-precision LES handle (nu_t held in FP32) are the
    handle's own nu_t at the slice indices, exactly."""
    from diablo3_amd.config import MIXED
    cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2, precision=MIXED)
    p = Problem(cfg)
    g = Solver(p)
    g.upload_state(p.initial_state(seed=2, theta_noise=1e-2))
    g.rem_div()
    g.step(time=0.0, time_step=1, create_new_flow=False)
    nut = g.download_les()[0]
    nxm, nym, nzm = 7, 12, 20
    _, sl = g.stats_movie(0.0, nxm, nym, nzm)
    ny = cfg.ny
    assert np.abs(nut).max() > 0
    assert np.array_equal(sl["nu_t_xz"], nut[1:ny + 1, nzm, :])
    assert np.array_equal(sl["nu_t_xy"], nut[nym])
    assert np.array_equal(sl["nu_t_yz"], nut[1:ny + 1, :, nxm])
    g.close()
