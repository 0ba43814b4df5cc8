"""The mean.h5 samples of save_stats_chan (for/diagnostics.f90:134-480, 1031-1037,
1235, 1838-1850) from the device statistics: the group names and the
profile each holds, as the reference writes them (its file names swap v and w:
group 'wme' holds vme, 'vme' holds wme, ...), through the series writers of
libdgpu_io.so (WriteStatH5_Y / WriteStatH5_X / WriteHDF5_real); and the movie
output of save_stats_chan(movie = .true.): the movie.h5 slices and the
mean_xz.h5 z-mean planes (WriteHDF5_XYplane / XZplane / ZYplane)."""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from . import flow_io

# (mean.h5 group, dgpu_stats profile), for/diagnostics.f90:140-235
PROFILE_GROUPS = [
    ("ume", "ume"), ("wme", "vme"), ("vme", "wme"), ("mke", "mke"),
    ("urms", "urms"), ("wrms", "vrms"), ("vrms", "wrms"), ("uw", "uv"), ("uv", "uw"), ("wv", "wv"),
    ("uu_dudx", "uu_dudx"), ("ww_dwdz", "vv_dvdy"), ("vu_dvdx", "wu_dwdx"), ("wu_dwdx", "vu_dvdx"),
    ("uw_dudz", "uv_dudy"), ("vw_dvdz", "wv_dwdy"), ("dudz", "dudy"), ("dvdz", "dwdy"),
    ("cp", "cp"), ("shear", "shear"),
    ("omega_x", "omega_x"), ("omega_z", "omega_y"), ("omega_y", "omega_z"),   # (224-235)
    ("epsilon", "epsilon"),      # compute_TKE_diss (1031)
    ("epsilon_m", "epsilon_m"),  # compute_MKE_diss (1235)
]
# per scalar n (for/diagnostics.f90:453-477): (group format, profile prefix)
SCALAR_GROUPS = [("thme{:02d}", "thme"), ("dthdz{:02d}", "dthdy"), ("thrms{:02d}", "thrms"),
                 ("thw{:02d}", "thv"), ("thw{:02d}_m", "thv_m"), ("pe_diss{:02d}", "pe_diss")]
ZSTAR_GROUPS = ["shear2_zstar", "epsilon_zstar", "TKE_zstar", "KE_zstar"]
# mean_xz.h5 planes in dgpu_stats_movie's order (for/diagnostics.f90:283-378,
# 756-767, 1076-1087, 1238-1245, 1338-1346, 1452-1460)
XZ_PLANES = ["ume_xz", "wme_xz", "vme_xz", "thme_xz", "uu_xz", "ww_xz", "vv_xz", "uw_xz", "wv_xz", "uv_xz",
             "epsilon_xz", "epsilon_m_xz", "thth_xz", "thu_xz", "thw_xz", "chi_xz"]


def movie_fields(n_th: int):
    """movie.h5 field names in dgpu_stats_movie's order: u1, u2 on GYF, u3, nu_t,
    epsilon, then the vorticity components omega_x, omega_y, omega_z under the
    reference's names (its omega_y section writes 'omegaZ', for/diagnostics.f90:
    1603-1660), then th1.. ("th", I0.1)."""
    return ["u", "w", "v", "nu_t", "epsilon", "omegaX", "omegaZ", "omegaY"] + [f"th{n + 1}" for n in range(n_th)]


def movie_indices(cfg, gyf: np.ndarray, xc: float, yc: float, zc: float):
    """init_chan_movie (for/domain.f90:380-401) at one rank: (NxMovie, NyMovie,
    NzMovie) from the slice positions; NyMovie = -1 when yc is outside
    [gyf(jstart), gyf(jend + 1)) (no x-z slice is written then)."""
    nxm = int(xc * cfg.nx / cfg.lx)
    nzm = int(zc * cfg.nz / cfg.lz)
    nzm = nzm - (nzm // cfg.nz) * cfg.nz
    nym = -1
    if gyf[cfg.jstart] <= yc < gyf[cfg.jend + 1]:
        i = 1
        while not (gyf[i] <= yc < gyf[i + 1]):
            i += 1
        nym = i
    return nxm, nym, nzm


def write_mean(fname: str, time: float, prof: Dict[str, np.ndarray], gyf: np.ndarray, ny: int, n_th: int,
               zstar: Optional[Dict[str, np.ndarray]] = None) -> None:
    """Append one save_stats sample to fname: 'time', 'gzf' and every profile
    (rows 1..Nyp), the scalar profiles, 'u1z_x0', and with zstar the BPE, the
    PDF, its first Nbin bin edges and the z*-binned means."""
    J = slice(1, ny + 1)
    flow_io.append_real(fname, "time", time)
    flow_io.append_profile(fname, "gzf", np.asarray(gyf)[J], time)
    for group, key in PROFILE_GROUPS:
        flow_io.append_profile(fname, group, prof[key][J], time)
    for n in range(n_th):
        for fmt, key in SCALAR_GROUPS:
            flow_io.append_profile(fname, fmt.format(n + 1), prof[f"{key}{n + 1:02d}"][J], time)
    flow_io.append_real(fname, "u1z_x0", prof["u1z_x0"])
    if "FTx_uu" in prof:
        flow_io.append_profile(fname, "FTx_uu", prof["FTx_uu"], time)
    if zstar is not None:
        nb = len(zstar["th1PDF"])
        for name in ZSTAR_GROUPS:
            flow_io.append_profile(fname, name, zstar[name], time)
        for n in range(n_th):
            flow_io.append_profile(fname, "thw_zstar", zstar[f"thw_zstar{n + 1:02d}"], time)
            flow_io.append_profile(fname, "chi_zstar", zstar[f"chi_zstar{n + 1:02d}"], time)
        flow_io.append_real(fname, "BPE", zstar["BPE"])
        flow_io.append_profile(fname, "th1PDF", zstar["th1PDF"], time)
        flow_io.append_profile(fname, "th1bin", zstar["th1bin"][:nb], time)


def save_stats(solver, fname: str, time: float) -> None:
    """save_stats_chan's mean.h5 sample of a Solver's state (device statistics)."""
    cfg = solver.cfg
    prof = solver.stats(time)
    z = solver.stats_zstar(time) if cfg.n_th >= 1 else None
    write_mean(fname, time, prof, solver.p.met["gyf"], cfg.ny, cfg.n_th, z)


def written_planes(cfg):
    """The mean_xz.h5 planes save_stats_chan(movie) writes for a configuration:
    the mean planes only when Nz > 1 and x is not homogeneous (756-767),
    epsilon_m_xz only when x is not homogeneous (compute_MKE_diss returns early,
    1117-1118), the fluctuation planes when Nz > 1, the scalar planes for scalar 1."""
    hx = cfg.ic_type not in (5, 6, 7, 8)  # homogeneousX, for/parameters.f90:282-287
    nz1 = cfg.nz > 1
    out = []
    if nz1 and not hx:
        out += ["ume_xz", "wme_xz", "vme_xz"] + (["thme_xz"] if cfg.n_th else [])
    if nz1:
        out += ["uu_xz", "ww_xz", "vv_xz", "uw_xz", "wv_xz", "uv_xz"]
    out.append("epsilon_xz")
    if not hx:
        out.append("epsilon_m_xz")
    if nz1 and cfg.n_th:
        out += ["thth_xz", "thu_xz", "thw_xz", "chi_xz"]
    return out


def write_movie(movie_fname: str, mean_xz_fname: str, time: float, planes: Dict[str, np.ndarray],
                slices: Dict[str, np.ndarray], cfg) -> None:
    """Append one movie sample: the slices to movie_fname and the z-mean planes to
    mean_xz_fname, under the conditions the reference writes them (Nz > 1, x not
    homogeneous, LES on, a scalar present, the x-z slice's row found)."""
    for name in written_planes(cfg):
        flow_io.append_plane(mean_xz_fname, name, planes[name], time)
    for f in movie_fields(cfg.n_th):
        if f == "nu_t" and not cfg.use_les:
            continue
        for cut in ("_xz", "_xy", "_yz"):
            if f + cut in slices:
                flow_io.append_plane(movie_fname, f + cut, slices[f + cut], time)


def save_movie(solver, time: float, nxm: int, nym: int, nzm: int, movie_fname: str = "movie.h5",
               mean_xz_fname: str = "mean_xz.h5") -> None:
    """save_stats_chan's movie output for a Solver's state (device statistics)."""
    planes, slices = solver.stats_movie(time, nxm, nym, nzm)
    write_movie(movie_fname, mean_xz_fname, time, planes, slices, solver.cfg)
