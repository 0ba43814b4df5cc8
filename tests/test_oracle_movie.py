"""The oracle's movie output of save_stats_chan (oracle/diagnostics.py: movie;
for/diagnostics.f90:283-378, 408-441, 485-617, 756-767, 1041-1087, 1238-1245,
1338-1346, 1452-1460, 1566-1727) pinned by what the mathematics fixes:

* each mean_xz plane is a z mean, so its x mean is the plane-averaged profile
  that save_stats_chan writes to mean.h5 (those profiles are pinned by closed
  forms in test_oracle_stats.py);
* th_1 = u1 makes thu = thth = uu (the scalar flux plane against its own
  variance);
* Taylor-Green cells u1 = b sin(ax) cos(bz), u3 = -a cos(ax) sin(bz) have the
  vertical vorticity (a^2 + b^2) sin(ax) sin(bz) (the reference's 'omegaZ'), and
  the u, v slices are the fields at the slice indices;
* the z mean of an epsilon z-y slice is a column of the epsilon_xz plane;
* init_chan_movie's indices (for/domain.f90:380-401) for positions on and
  between grid lines."""
import numpy as np

import oracle
from diablo3_amd import configs, stats_io
from diablo3_amd.ic import FourierState
from diablo3_amd.problem import Problem
from oracle import diagnostics as od


def front(nx=16, ny=17, nz=16, seed=3):
    cfg = configs.front_les(nx, ny, nz, 5000.0, 2e-4, 1e-2)
    cfg.use_les = False
    p = Problem(cfg)
    o = oracle.Oracle(p)
    o.upload_state(p.initial_state(seed=seed, theta_noise=1e-2))
    o.rem_div()
    o.step()
    return cfg, p, o


def test_planes_average_to_the_profiles():
    cfg, p, o = front()
    ny, nx = cfg.ny, cfg.nx
    prof = od.save_stats(o, 0.0)
    planes, _ = od.movie(o, 3, 5, 2)
    J = slice(1, ny + 1)
    G = slice(1, ny)   # GY rows 2..Nyp of the (1..Nyp) plane rows

    def close(a, b, scale=None):
        return np.abs(a - b).max() <= 1e-12 * (np.abs(b).max() if scale is None else scale)
    xm = {k: v.mean(axis=1) for k, v in planes.items()}
    assert close(xm["uu_xz"], prof["urms"][J] ** 2)
    assert close(xm["ww_xz"], prof["vrms"][J] ** 2)
    assert close(xm["vv_xz"], prof["wrms"][J] ** 2)
    assert close(xm["uw_xz"][G], prof["uv"][2:ny + 1]) and not planes["uw_xz"][0].any()
    assert close(xm["wv_xz"][G], prof["wv"][2:ny + 1])
    assert close(xm["uv_xz"], prof["uw"][J])
    assert close(xm["epsilon_xz"], prof["epsilon"][J])
    assert close(xm["epsilon_m_xz"], prof["epsilon_m"][J])
    assert close(xm["thth_xz"], prof["thrms01"][J] ** 2)
    assert close(xm["thw_xz"], prof["thv01"][J])
    assert close(xm["chi_xz"], prof["pe_diss01"][J])
    assert close(xm["ume_xz"], prof["ume"][J]) and close(xm["vme_xz"], prof["wme"][J])
    assert close(xm["thme_xz"], prof["thme01"][J])
    # wme_xz is v on GYF: the mean of the two GY rows around it
    # (the mean of v is at round-off here: compared on the scale of v')
    assert close(xm["wme_xz"], 0.5 * (prof["vme"][1:ny + 1] + prof["vme"][2:ny + 2]), prof["vrms"].max())


def test_scalar_flux_plane_of_a_scalar_equal_to_u1():
    cfg, p, o = front(seed=5)
    s = o.download_state()
    o.upload_state(FourierState(s.cu1, s.cu2, s.cu3, s.cp, [s.cu1.copy()]))
    planes, _ = od.movie(o, 0, 4, 0)
    # away from the wall rows, where the BC posts of u1 and th differ
    R = slice(1, cfg.ny - 1)
    scale = np.abs(planes["uu_xz"][R]).max()
    assert np.abs(planes["thu_xz"][R] - planes["uu_xz"][R]).max() <= 1e-13 * scale
    assert np.abs(planes["thth_xz"][R] - planes["uu_xz"][R]).max() <= 1e-13 * scale


def test_taylor_green_slices():
    cfg = configs.small_dns(16, 17, 16)
    cfg.n_th = 0
    cfg.ri, cfg.pr = [], []
    p = Problem(cfg)
    o = oracle.Oracle(p)
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    X = 2 * np.pi * (np.arange(cfg.nx) * cfg.lx / cfg.nx)[None, None, :] / cfg.lx
    Z = 2 * np.pi * (np.arange(cfg.nz) * cfg.lz / cfg.nz)[None, :, None] / cfg.lz
    a, b = p.kx[1], p.kz[1]
    u1 = np.broadcast_to(b * np.sin(X) * np.cos(Z), shp).copy()
    u3 = np.broadcast_to(-a * np.cos(X) * np.sin(Z), shp).copy()
    zero = np.zeros(shp)
    four = o.fft_to_fourier
    o.upload_state(FourierState(four(u1), four(zero), four(u3), four(zero), []))
    nxm, nym, nzm = 5, 7, 3
    _, sl = od.movie(o, nxm, nym, nzm)
    ny = cfg.ny
    oy = (a * a + b * b) * np.sin(X) * np.sin(Z)
    oy = np.broadcast_to(oy, shp)
    assert np.abs(sl["omegaZ_xy"] - oy[nym]).max() <= 1e-13 * (a * a + b * b)
    # rows 2..Nyp (the no-slip BC post zeroes u1 and u3 on row 1)
    assert np.abs(sl["omegaZ_xz"][1:] - oy[2:ny + 1, nzm, :]).max() <= 1e-13 * (a * a + b * b)
    assert np.abs(sl["omegaZ_yz"][1:] - oy[2:ny + 1, :, nxm]).max() <= 1e-13 * (a * a + b * b)
    assert np.abs(sl["u_yz"][1:] - u1[2:ny + 1, :, nxm]).max() <= 1e-14 * b
    assert not sl["u_xz"][0].any()
    assert np.abs(sl["v_xy"] - u3[nym]).max() <= 1e-14 * a
    assert not sl["w_xz"].any() and sl["u_xz"].shape == (ny, cfg.nx) and sl["u_yz"].shape == (ny, cfg.nz)


def test_epsilon_slice_against_its_plane():
    cfg, p, o = front(seed=7)
    nxm = 6
    planes, sl = od.movie(o, nxm, 8, 1)
    col = sl["epsilon_yz"].mean(axis=1)
    assert np.abs(col - planes["epsilon_xz"][:, nxm]).max() <= 1e-13 * np.abs(col).max()
    assert not sl["epsilon_xz"][0].any()   # row 1: epsilon lives on GY rows 2..Nyp


def test_movie_indices():
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    gyf = p.met["gyf"]
    nxm, nym, nzm = stats_io.movie_indices(cfg, gyf, 0.5 * cfg.lx, gyf[6], 0.26 * cfg.lz)
    assert (nxm, nym, nzm) == (cfg.nx // 2, 6, int(0.26 * cfg.nz))
    assert stats_io.movie_indices(cfg, gyf, 0.0, 0.5 * (gyf[6] + gyf[7]), 0.0)[1] == 6
    assert stats_io.movie_indices(cfg, gyf, 0.0, gyf[cfg.jstart] - 1e-9, 0.0)[1] == -1
