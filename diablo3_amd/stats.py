"""Names of the profiles dgpu_stats returns, in its output order (the code
variable names of save_stats_chan, for/diagnostics.f90:3-669)."""
PROFILES = ["ume", "vme", "wme", "mke", "urms", "vrms", "wrms", "uw", "uv", "wv",
            "uu_dudx", "vv_dvdy", "wu_dwdx", "vu_dvdx", "uv_dudy", "wv_dwdy",
            "dudy", "dwdy", "shear", "epsilon", "epsilon_m", "cp", "u1y_left",
            "omega_x", "omega_y", "omega_z"]
PER_SCALAR = ["thme", "thrms", "thv", "thv_m", "dthdy", "pe_diss"]
