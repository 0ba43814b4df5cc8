"""The mean.h5 samples of save_stats_chan (for/diagnostics.f90:134-480, 1031-1037,
1235, 1838-1850) from the device statistics: the group names and the
profile each holds, as the reference writes them (its file names swap v and w:
group 'wme' holds vme, 'vme' holds wme, ...), through the series writers of
libdgpu_io.so (WriteStatH5_Y / WriteStatH5_X / WriteHDF5_real)."""
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
