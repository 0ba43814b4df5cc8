"""WriteHDF5 / ReadHDF5 with the device state (dgpu_write_flow, dgpu_read_flow)
against the oracle's transforms and the numpy restatement of swapzy / g2gf /
gf2g (oracle/flow_io.py), the file read back through the HDF5 C library itself.

Bars: each dataset within 1e-13 (relative max) of the oracle's physical field,
the writer's post-write state (the reference's g2gf / gf2g and transform round
trip) within 1e-13 of the oracle's, and a fresh handle that reads the file holds
the writer's post-write state on the rows the file carries, to 1e-15."""
import numpy as np
import pytest

import h5raw
import oracle
from diablo3_amd import configs, layout
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver
from oracle import flow_io as ofi

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5raw.available(), reason="HDF5 C library not in the image")]


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_checkpoint_write_read(tmp_path):
    from diablo3_amd import flow_io
    cfg = configs.small_dns(32, 33, 32)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    p = Problem(cfg)
    g = Solver(p)
    s0 = p.initial_state(seed=9, theta_noise=1e-3)
    g.upload_state(s0)
    g.rem_div()
    for n in range(2):
        g.step(time=n * cfg.delta_t, time_step=n + 1, create_new_flow=True)
    s = g.download_state()
    o = oracle.Oracle(p, threads=16)

    def phys(four):  # the oracle's physical field in the Fortran physical layout
        return layout.to_fortran_physical(cfg, o.fft_to_physical(four))

    path = str(tmp_path / "out.h5")
    times = {"time": 2 * cfg.delta_t, "save_flow_time": 1.0, "save_stats_time": 0.5, "save_movie_time": 0.25,
             "time_step": 2}
    print(f"write_flow {nx}x{ny}x{nz}", flush=True)
    flow_io.write_flow(g, path, times, save_pressure=True)

    u1, u2, u3, pp = phys(s.cu1), phys(s.cu2), phys(s.cu3), phys(s.cp)
    th = phys(s.cth[0])
    u2f = ofi.g2gf(u2, nx, ny, nz)
    h = h5raw.File(path)
    try:
        assert list(h.attr("/", "Resolution", "i32")[0]) == [nx, ny, nz]
        assert h.attr("/Timestep", "Time_Step", "i32")[0][0] == 2
        for name, src in (("U", u1), ("W", u2f), ("V", u3), ("TH1", th), ("P", pp)):
            data, chunk, _ = h.dataset("/Timestep/" + name)
            err = relmax(data, ofi.swapzy(src, nx, ny, nz))
            print(name, f"{err:.2e}", flush=True)
            assert data.shape == (nz, ny, nx) and chunk == (nz, 1, nx) and err <= 1e-13, name
    finally:
        h.close()

    # the writer's state afterwards: u2 through g2gf then gf2g, all fields through
    # the transform round trip (for/phdf5.f90:1112-1114, 1171, 1187-1192)
    s2 = g.download_state()
    cp_back = layout.from_fortran_physical
    want = {"cu1": o.fft_to_fourier(cp_back(cfg, u1)),
            "cu2": o.fft_to_fourier(cp_back(cfg, ofi.gf2g(u2f, nx, ny, nz))),
            "cu3": o.fft_to_fourier(cp_back(cfg, u3)), "cp": o.fft_to_fourier(cp_back(cfg, pp)),
            "cth0": o.fft_to_fourier(cp_back(cfg, th))}
    got = {"cu1": s2.cu1, "cu2": s2.cu2, "cu3": s2.cu3, "cp": s2.cp, "cth0": s2.cth[0]}
    for k in want:
        err = relmax(got[k], want[k])
        print("post-write", k, f"{err:.2e}", flush=True)
        assert err <= 1e-13, k

    # a fresh handle reads the file
    g2 = Solver(p)
    t, had_p = flow_io.read_flow(g2, path)
    assert had_p and t == times
    s3 = g2.download_state()
    for k, a, b in (("cu1", s3.cu1, s2.cu1), ("cu3", s3.cu3, s2.cu3), ("cp", s3.cp, s2.cp),
                    ("cth0", s3.cth[0], s2.cth[0])):
        assert relmax(a[1:ny + 1], b[1:ny + 1]) <= 1e-15, k
    assert relmax(s3.cu2[1:ny + 2], s2.cu2[1:ny + 2]) <= 1e-15
    g.close()
    g2.close()


def test_read_flow_with_variable_dt_runs_courant(tmp_path):
    """ReadHDF5 calls courant on the velocity it has read when variable_dt is set
    (for/phdf5.f90:1437-1439): after dgpu_read_flow the next dgpu_courant returns
    courant's delta_t for that velocity, as the oracle's courant of the state read
    (limits, caps and event scheduling) gives it; without variable_dt it has not run."""
    from diablo3_amd import flow_io
    cfg = configs.small_dns(32, 33, 32)
    p = Problem(cfg)
    g = Solver(p)
    s0 = p.initial_state(seed=4, theta_noise=1e-3)
    for a in (s0.cu1, s0.cu2, s0.cu3):
        a *= 500.0  # velocities large enough that the CFL limit is the smallest
    g.upload_state(s0)
    g.rem_div()
    g.step(time=0.0, time_step=1, create_new_flow=True)
    path = str(tmp_path / "restart.h5")
    times = {"time": 0.5, "save_flow_time": 0.75, "save_stats_time": 0.6, "save_movie_time": 0.25,
             "time_step": 1}
    print("write_flow, then read_flow with and without variable_dt", flush=True)
    flow_io.write_flow(g, path, times, save_pressure=True)
    g.close()
    g0 = Solver(p)
    flow_io.read_flow(g0, path)
    assert g0.courant(0.5, 10.0, 0.6, 0.75) == (0.0, False)
    g0.close()
    g2 = Solver(p)
    cfl = 0.5
    g2.set_variable_dt(True, 1, cfl)
    t, _ = flow_io.read_flow(g2, path)
    dt, ran = g2.courant(t["time"], 10.0, t["save_stats_time"], t["save_flow_time"])
    o = oracle.Oracle(p, threads=16)
    o.upload_state(g2.download_state())
    lim = o.courant_dt_state(cfl)
    want = oracle.lib().oracle_courant_schedule(lim, t["time"], 10.0, t["save_stats_time"], t["save_flow_time"])
    print(f"courant after read: device dt {dt:.15e}, oracle {want:.15e} (limit {lim:.6e})", flush=True)
    assert ran and abs(dt - want) <= 1e-12 * want
    assert lim < 0.5 * (0.5 * (min(cfg.lx / cfg.nx, cfg.lz / cfg.nz) ** 2) / cfg.nu)  # the CFL limit is active
    g2.close()


def test_save_stats_mean_h5(tmp_path):
    """save_stats_chan's mean.h5 sample from the device statistics: the groups
    hold the device profiles under the reference's names (a second evaluation of
    the same state gives the same profiles; the z* sums to rounding)."""
    from diablo3_amd import stats_io
    cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
    cfg.use_les = False
    ny = cfg.ny
    p = Problem(cfg)
    g = Solver(p)
    g.upload_state(p.initial_state(seed=3, theta_noise=1e-2))
    g.rem_div()
    g.step(time=0.0, time_step=1, create_new_flow=True)
    path = str(tmp_path / "mean.h5")
    print("save_stats -> mean.h5", flush=True)
    stats_io.save_stats(g, path, 0.25)
    prof, z = g.stats(0.25), g.stats_zstar(0.25)
    h = h5raw.File(path)
    try:
        J = slice(1, ny + 1)
        for group, key in stats_io.PROFILE_GROUPS:
            assert np.array_equal(h.dataset(f"/{group}/0001")[0], prof[key][J]), group
        assert np.array_equal(h.dataset("/thw01_m/0001")[0], prof["thv_m01"][J])
        assert abs(h.dataset("/BPE/0001")[0][0] - z["BPE"]) <= 1e-13 * abs(z["BPE"])
        for name in ("th1PDF", "KE_zstar", "chi_zstar"):
            a = h.dataset(f"/{name}/0001")[0]
            b = z["chi_zstar01"] if name == "chi_zstar" else z[name]
            m = ~np.isnan(b)
            assert np.array_equal(np.isnan(a), ~m) and np.allclose(a[m], b[m], rtol=1e-12, atol=0), name
        assert h.attr("/time", "SAMPLES", "i32")[0][0] == 1
    finally:
        h.close()
    g.close()


def _front_state(seed):
    cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
    cfg.use_les = False
    p = Problem(cfg)
    g = Solver(p)
    g.upload_state(p.initial_state(seed=seed, theta_noise=1e-2))
    g.rem_div()
    g.step(time=0.0, time_step=1, create_new_flow=True)
    return cfg, p, g


def test_poisson_p_matches_oracle():
    """dgpu_poisson_p (poisson_p_chan, for/solver.f90:1382-1560) against the
    oracle's spectral-product restatement, relative max 1e-11 on rows 1..Nyp; the
    velocity and scalars are not touched."""
    cfg, p, g = _front_state(4)
    ny = cfg.ny
    s = g.download_state()
    o = oracle.Oracle(p, threads=16)
    o.upload_state(s)
    want = ofi.poisson_p(o)
    print("poisson_p 32x33x32", flush=True)
    g.poisson_p()
    s2 = g.download_state()
    err = relmax(s2.cp[1:ny + 1], want[1:ny + 1])
    print(f"cp {err:.2e}", flush=True)
    assert err <= 1e-11
    for a, b in ((s2.cu1, s.cu1), (s2.cu2, s.cu2), (s2.cu3, s.cu3), (s2.cth[0], s.cth[0])):
        assert np.array_equal(a, b)
    g.close()


def test_read_flow_without_pressure_recomputes_it(tmp_path):
    """ReadHDF5 of a file without P calls poisson_p_chan (for/phdf5.f90:1451-1454):
    the pressure of the fresh handle is the oracle's poisson_p of the state read."""
    from diablo3_amd import flow_io
    cfg, p, g = _front_state(6)
    ny = cfg.ny
    path = str(tmp_path / "nop.h5")
    flow_io.write_flow(g, path, {"time": 0.5, "time_step": 3}, save_pressure=False)
    g2 = Solver(p)
    print("read_flow without P", flush=True)
    _, had_p = flow_io.read_flow(g2, path)
    assert not had_p
    s3 = g2.download_state()
    o = oracle.Oracle(p, threads=16)
    o.upload_state(s3)
    want = ofi.poisson_p(o)
    err = relmax(s3.cp[1:ny + 1], want[1:ny + 1])
    print(f"cp {err:.2e}", flush=True)
    assert err <= 1e-11 and np.abs(s3.cp[1:ny + 1]).max() > 0
    g.close()
    g2.close()


def test_save_movie_files(tmp_path):
    """save_stats_chan's movie output through the device and the plane writers:
    movie.h5 and mean_xz.h5 hold the arrays dgpu_stats_movie returns."""
    from diablo3_amd import stats_io
    cfg, p, g = _front_state(8)
    nxm, nym, nzm = stats_io.movie_indices(cfg, p.met["gyf"], 0.5 * cfg.lx, 0.0, 0.5 * cfg.lz)
    mv, mx = str(tmp_path / "movie.h5"), str(tmp_path / "mean_xz.h5")
    print("save_movie", nxm, nym, nzm, flush=True)
    stats_io.save_movie(g, 0.5, nxm, nym, nzm, mv, mx)
    planes, slices = g.stats_movie(0.5, nxm, nym, nzm)
    h = h5raw.File(mv)
    try:
        for name in ("u_xz", "w_xy", "th1_yz", "epsilon_xz", "omegaX_xy"):
            d = h.dataset(f"/{name}/0001")[0]
            assert np.allclose(d, slices[name], rtol=1e-13, atol=1e-300), name
        assert h.attr("/v_yz/0001", "Time", "f64")[0][0] == 0.5
    finally:
        h.close()
    h = h5raw.File(mx)
    try:
        for name in stats_io.written_planes(cfg):
            d = h.dataset(f"/{name}/0001")[0]
            assert d.shape == (cfg.ny, cfg.nx) and np.allclose(d, planes[name], rtol=1e-13, atol=1e-18), name
    finally:
        h.close()
    g.close()
