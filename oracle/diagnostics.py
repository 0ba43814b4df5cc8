"""CPU restatement of the energy-budget statistics of save_stats_chan
(for/diagnostics.f90:3-669 with compute_averages 676-799, compute_TKE_diss
804-1092, compute_MKE_diss 1097-1249, compute_TKE 1254-1360,
compute_TKE_Production 1366-1519).

TEST INFRASTRUCTURE ONLY: the checker for the device statistics.  The BC posts and
the 2D transforms are the C oracle's; the loops of the reference are restated
here with numpy over whole planes (sums in a different order: round-off only).

Profiles are indexed j = 0..Nyp+1 like the reference's (0:Nyp+1) arrays; entries
the reference never sets are 0 (its module arrays start at zero).  Not restated:
movie planes, z* binning, the BPE PDF and the FTx_uu spectrum (out of scope), the
vorticity rms, and the closing fft_xz_to_physical/fft_xz_to_fourier round trip of
the state (a round-off-level change).
"""
import numpy as np

# profile order of dgpu_stats (include/diablo_gpu.h)
BASE = ["ume", "vme", "wme", "mke", "urms", "vrms", "wrms", "uw", "uv", "wv",
        "uu_dudx", "vv_dvdy", "wu_dwdx", "vu_dvdx", "uv_dudy", "wv_dwdy",
        "dudy", "dwdy", "shear", "epsilon", "epsilon_m", "cp", "u1y_left",
        "omega_x", "omega_y", "omega_z"]
PER_TH = ["thme", "thrms", "thv", "thv_m", "dthdy", "pe_diss"]


def names(n_th):
    return BASE + [f"{b}{n + 1:02d}" for n in range(n_th) for b in PER_TH]


def homogeneous_x(cfg):
    """for/parameters.f90:282-287"""
    return cfg.ic_type not in (5, 6, 7, 8)


def save_stats(o, time=0.0, points=False):
    """Statistics of the oracle's current state; applies the BC posts to it, as
    save_stats_chan does (for/diagnostics.f90:26-27).  Returns {name: profile}
    plus 'u1z_x0' (a scalar); with points=True also '_points', the per-point fields
    the reference bins on z* surfaces ([j][z][x] arrays) and th of scalar 1."""
    p = o.p
    cfg, met = p.cfg, p.met
    ny, nx, nz = cfg.ny, cfg.nx, cfg.nz
    n2 = ny + 2
    dy, dyf, gy, gyf = met["dy"], met["dyf"], met["gy"], met["gyf"]
    kx = p.kx[:cfg.kxr][None, None, :]
    kz = p.kz[None, :, None]
    hx = homogeneous_x(cfg)
    nu = cfg.nu
    twf = (1.0 / (cfg.ro_inv / cfg.delta)) * cfg.dthdx[0] if cfg.n_th else 0.0  # thermal-wind shear
    o.bc_post(time)
    s = o.download_state()
    cr1, cr2, cr3, cp = s.cu1, s.cu2, s.cu3, s.cp
    phys = o.fft_to_physical
    u1, u2, u3 = phys(cr1), phys(cr2), phys(cr3)
    th = [phys(t) for t in s.cth]
    out = {k: np.zeros(n2) for k in names(cfg.n_th)}
    NxNz = float(nx * nz)

    # ---- compute_averages (676-799)
    ume, vme, wme = cr1[:, 0, 0].real.copy(), cr2[:, 0, 0].real.copy(), cr3[:, 0, 0].real.copy()
    thme = [t[:, 0, 0].real.copy() for t in s.cth]
    # z-line means [j][x]; ume_xy etc. on j = 1..Nyp, vme_xy on 1..Nyp+1 (row 0 unused)
    ume_xy = np.zeros((n2, nx)); wme_xy = np.zeros((n2, nx)); vme_xy = np.zeros((n2, nx))
    thme_xy = [np.zeros((n2, nx)) for _ in th]
    if nz > 1 and not hx:
        ume_xy[1:ny + 1] = u1[1:ny + 1].sum(axis=1) / nz
        wme_xy[1:ny + 1] = u3[1:ny + 1].sum(axis=1) / nz
        vme_xy[1:ny + 2] = u2[1:ny + 2].sum(axis=1) / nz
        for n, t in enumerate(th):
            thme_xy[n][1:ny + 1] = t[1:ny + 1].sum(axis=1) / nz
    else:
        ume_xy[1:ny + 1] = ume[1:ny + 1, None]
        wme_xy[1:ny + 1] = wme[1:ny + 1, None]
        vme_xy[1:ny + 2] = vme[1:ny + 2, None]
        for n in range(len(th)):
            thme_xy[n][1:ny + 1] = thme[n][1:ny + 1, None]
    out["ume"], out["vme"], out["wme"] = ume, vme, wme
    out["cp"] = cp[:, 0, 0].real.copy()

    # ---- compute_TKE_diss (804-1092)
    def d_phys(c, k0=0, jl=1, jh=ny):
        a = np.zeros_like(c)
        a[jl:jh + 1, k0:] = c[jl:jh + 1, k0:]
        return phys(a)
    k_start = 0 if hx else 1
    interp = lambda f, j: (dyf[j - 1] * f[j] ** 2 + dyf[j] * f[j - 1] ** 2) / (2.0 * dy[j])
    eps = np.zeros(n2)
    pts = {"epsilon_zstar": np.zeros_like(u1), "shear2_zstar": np.zeros_like(u1)}
    a1 = d_phys(1j * kx * cr1, k_start)
    a2 = d_phys(1j * kx * cr2, k_start, 2, ny)
    a4 = d_phys(1j * kx * cr3, k_start)
    a5 = d_phys(1j * kz * cr1)
    a8 = d_phys(1j * kz * cr2, 0, 2, ny)
    a9 = d_phys(1j * kz * cr3)
    for j in range(2, ny + 1):
        t = interp(a1, j) + a2[j] ** 2
        t = t + (((u1[j] - ume_xy[j][None, :]) - (u1[j - 1] - ume_xy[j - 1][None, :])) / dy[j]) ** 2
        t = t + interp(a4, j) + interp(a5, j)
        t = t + (((u2[j + 1] - vme_xy[j + 1][None, :]) - (u2[j - 1] - vme_xy[j - 1][None, :]))
                 / (gy[j + 1] - gy[j - 1])) ** 2
        t = t + (((u3[j] - wme_xy[j][None, :]) - (u3[j - 1] - wme_xy[j - 1][None, :])) / dy[j]) ** 2
        t = t + a8[j] ** 2 + interp(a9, j)
        eps[j] = t.sum()
        pts["epsilon_zstar"][j] = nu * t       # f1 = nu * f1 (for/diagnostics.f90:1021)
    out["epsilon"] = nu * eps / NxNz

    # ---- compute_MKE_diss (1097-1249); gradients of the z-mean fields
    dudx_m = np.zeros((n2, nx)); dvdx_m = np.zeros((n2, nx)); dwdx_m = np.zeros((n2, nx))
    epm = np.zeros(n2)
    epm_xy = np.zeros((n2, nx))
    if not hx:
        def mean_dx(c):
            a = np.zeros_like(c)
            a[1:ny + 1, 0] = 1j * kx[0, 0] * c[1:ny + 1, 0]
            return phys(a)[:, 0, :]          # s1(i, 0, j): the k = 0 row, any z
        m1, m2, m4 = mean_dx(cr1), mean_dx(cr2), mean_dx(cr3)
        for j in range(1, ny + 1):
            t = (dyf[j - 1] * m1[j] ** 2 + dyf[j] * m1[j - 1] ** 2) / (2.0 * dy[j])
            t = t + m2[j] ** 2
            t = t + (dyf[j - 1] * m4[j] ** 2 + dyf[j] * m4[j - 1] ** 2) / (2.0 * dy[j])
            if j >= 2:
                t = t + ((ume_xy[j] - ume_xy[j - 1]) / dy[j]) ** 2
                t = t + ((vme_xy[j + 1] - vme_xy[j - 1]) / (gy[j + 1] - gy[j - 1])) ** 2
                t = t + ((wme_xy[j] - wme_xy[j - 1]) / dy[j] + twf) ** 2
            epm[j] = t.sum()
            epm_xy[j] = t
        dudx_m[1:ny + 1], dvdx_m[1:ny + 1], dwdx_m[1:ny + 1] = m1[1:ny + 1], m2[1:ny + 1], m4[1:ny + 1]
    out["epsilon_m"] = nu * epm / float(nx)

    # ---- compute_TKE (1254-1360)
    f1 = u1 - ume_xy[:, None, :]
    f2 = u2 - vme_xy[:, None, :]
    f3 = u3 - wme_xy[:, None, :]
    J = slice(1, ny + 1)
    uu_xy = np.zeros((n2, nx)); vv_xy = np.zeros((n2, nx)); ww_xy = np.zeros((n2, nx))
    pts["TKE_zstar"] = np.zeros_like(u1)
    pts["TKE_zstar"][J] = 0.5 * f1[J] ** 2 + 0.5 * f2[J] ** 2 + 0.5 * f3[J] ** 2   # (1273-1275)
    pts["KE_zstar"] = 0.5 * (u1 ** 2 + u2 ** 2 + u3 ** 2)                          # (1354)
    for name, f, xy in (("urms", f1, uu_xy), ("vrms", f2, vv_xy), ("wrms", f3, ww_xy)):
        out[name][J] = np.sqrt((f[J] ** 2).sum(axis=(1, 2)) / NxNz)
        xy[J] = (f[J] ** 2).sum(axis=1) / nz

    # ---- compute_TKE_Production (1366-1519)
    out["uw"][J] = (f1[J] * f3[J]).sum(axis=(1, 2)) / NxNz
    vvar_xy = np.zeros((n2, nx)); uvar_xy = np.zeros((n2, nx)); wvar_xy = np.zeros((n2, nx))
    vvar_xy[J] = (f1[J] * f3[J]).sum(axis=1) / nz
    for j in range(2, ny + 1):
        iu = (dyf[j - 1] * f1[j] + dyf[j] * f1[j - 1]) / (2.0 * dy[j])
        iw = (dyf[j - 1] * f3[j] + dyf[j] * f3[j - 1]) / (2.0 * dy[j])
        out["uv"][j] = (iu * f2[j]).sum() / NxNz
        out["wv"][j] = (iw * f2[j]).sum() / NxNz
        uvar_xy[j] = (iu * f2[j]).sum(axis=0) / nz
        wvar_xy[j] = (iw * f2[j]).sum(axis=0) / nz
    if nz > 1:
        for j in range(1, ny + 1):
            out["uu_dudx"][j] = (uu_xy[j] * dudx_m[j]).sum() / float(nx)
            out["wu_dwdx"][j] = (vvar_xy[j] * dwdx_m[j]).sum() / float(nx)
        for j in range(2, ny + 1):
            out["vu_dvdx"][j] = (uvar_xy[j] * dvdx_m[j]).sum() / float(nx)
            out["vv_dvdy"][j] = ((vme_xy[j + 1] - vme_xy[j - 1]) / (gy[j + 1] - gy[j - 1]) * vv_xy[j]).sum() / float(nx)
            out["uv_dudy"][j] = ((ume_xy[j] - ume_xy[j - 1]) / dy[j] * uvar_xy[j]).sum() / float(nx)
            out["wv_dwdy"][j] = (((wme_xy[j] - wme_xy[j - 1]) / dy[j] + twf) * wvar_xy[j]).sum() / float(nx)

    # ---- save_stats_chan body (44-120): mke, dudy, dwdy, shear
    for j in range(2, ny + 1):
        t = (dyf[j - 1] * ume_xy[j] ** 2 + dyf[j] * ume_xy[j - 1] ** 2) / (2.0 * dy[j])
        t = t + (dyf[j - 1] * (wme_xy[j] + twf * (gyf[j] - 0.5 * cfg.ly)) ** 2 +
                 dyf[j] * (wme_xy[j - 1] + twf * (gyf[j - 1] - 0.5 * cfg.ly)) ** 2) / (2.0 * dy[j])
        t = t + vme_xy[j] ** 2
        out["mke"][j] = nz * t.sum()
    out["mke"] = 0.5 * out["mke"] / NxNz
    for j in range(1, ny + 1):
        out["dudy"][j] = (ume[j] - ume[j - 1]) / (gyf[j] - gyf[j - 1])
        out["dwdy"][j] = (wme[j] - wme[j - 1]) / (gyf[j] - gyf[j - 1])
        sh = ((u1[j + 1] - u1[j - 1]) / (2.0 * dyf[j])) ** 2 + ((u3[j + 1] - u3[j - 1]) / (2.0 * dyf[j])) ** 2
        out["shear"][j] = sh.sum() / NxNz
        pts["shear2_zstar"][j] = sh

    # ---- compute_Vorticity (1525-1727): rms of the spectral curl, omega_x less
    # the thermal-wind shear
    def curl_rms(build, shift=0.0):
        a = np.zeros_like(cr1)
        for j in range(1, ny + 1):
            a[j] = build(j)
        f = phys(a)
        r = np.zeros(n2)
        for j in range(1, ny + 1):
            r[j] = np.sqrt(((f[j] - shift) ** 2).sum() / NxNz)
        return r, f - shift
    out["omega_x"], pts["omega_x"] = curl_rms(lambda j: 1j * kz * 0.5 * (cr2[j + 1] + cr2[j])
                                              - (cr3[j + 1] - cr3[j - 1]) / (2.0 * dyf[j]), twf)
    out["omega_y"], pts["omega_y"] = curl_rms(lambda j: 1j * kx * cr3[j] - 1j * kz * cr1[j])
    out["omega_z"], pts["omega_z"] = curl_rms(lambda j: (cr1[j + 1] - cr1[j - 1]) / (2.0 * dyf[j])
                                              - 1j * kx * 0.5 * (cr2[j + 1] + cr2[j]))

    # ---- scalars (140-410)
    gxv = np.array([(i * cfg.lx) / nx for i in range(nx)])
    for n, t in enumerate(th):
        tag = f"{n + 1:02d}"
        c = s.cth[n]
        r1 = d_phys(1j * kx * c)
        a = np.zeros_like(c)
        a[1:ny + 1] = (s.cth[0][2:ny + 2] - s.cth[0][0:ny]) / (gyf[2:ny + 2] - gyf[0:ny])[:, None, None]
        r2 = phys(a)
        r3 = d_phys(1j * kz * c)
        ft = t - thme_xy[n][:, None, :]
        out["thrms" + tag][J] = np.sqrt((ft[J] ** 2).sum(axis=(1, 2)) / NxNz)
        pts["thw_zstar" + tag] = np.zeros_like(u1)
        pts["chi_zstar" + tag] = np.zeros_like(u1)
        for j in range(1, ny + 1):
            w = 0.5 * (u2[j] + u2[j + 1]) - 0.5 * (vme_xy[j] + vme_xy[j + 1])[None, :]
            out["thv" + tag][j] = (ft[j] * w).sum() / NxNz
            pts["thw_zstar" + tag][j] = ft[j] * w                                    # (305-307)
            out["thv_m" + tag][j] = ((thme_xy[n][j] + cfg.dthdx[n] * (gxv - 0.5 * cfg.lx)) *
                                     (0.5 * (vme_xy[j] + vme_xy[j + 1]))).sum() / float(nx)
            out["dthdy" + tag][j] = (thme[n][j] - thme[n][j - 1]) / (gyf[j] - gyf[j - 1])
            g2 = (r1[j] + cfg.dthdx[n]) ** 2 + r2[j] ** 2 + (r3[j] + cfg.dthdz[n]) ** 2
            out["pe_diss" + tag][j] = g2.sum() / NxNz
            pts["chi_zstar" + tag][j] = g2                                           # (359-361)
        out["thme" + tag] = thme[n]
    # <u1 y> at the x = 0 edge, integrated in y (for/diagnostics.f90:388-402, 1964-1980)
    u1y = np.zeros(n2)
    u1y[1:ny + 1] = gyf[1:ny + 1] * u1[1:ny + 1, :, 0].sum(axis=1) / float(nz)
    out["u1y_left"] = u1y
    res = 0.0
    for j in range(2, ny + 1):
        res = res + 0.5 * (u1y[j] + u1y[j - 1]) * dy[j]
    out["u1z_x0"] = res / cfg.ly
    # FTx_uu (641-663): kz = 0 row of u1, |c|^2 integrated in y by trapezoids
    cuu = np.zeros((n2, nx // 2))
    cuu[:, :cfg.kxr] = (s.cu1[:, 0, :] * np.conj(s.cu1[:, 0, :])).real
    ftx = np.zeros(nx // 2)
    for j in range(2, ny + 1):
        ftx = ftx + 0.5 * (cuu[j] + cuu[j - 1]) * dy[j]
    out["FTx_uu"] = ftx / cfg.ly
    if points:
        pts["th0"] = th[0] if th else None
        out["_points"] = pts
        # mean_xz.h5 planes (283-378, 756-767, 1076-1087, 1238-1245, 1338-1346,
        # 1452-1460), rows 1..Nyp
        pl = {"ume_xz": ume_xy, "wme_xz": np.zeros((n2, nx)), "vme_xz": wme_xy,
              "uu_xz": uu_xy, "ww_xz": vv_xy, "vv_xz": ww_xy, "uw_xz": uvar_xy, "wv_xz": wvar_xy,
              "uv_xz": vvar_xy, "epsilon_xz": pts["epsilon_zstar"].sum(axis=1) / nz, "epsilon_m_xz": nu * epm_xy}
        for j in range(1, ny + 1):
            pl["wme_xz"][j] = (0.5 * (u2[j] + u2[j + 1])).sum(axis=0) / nz   # var_xy (723)
        if th:
            ft = th[0] - thme_xy[0][:, None, :]
            pl["thme_xz"] = thme_xy[0]
            pl["thth_xz"] = (ft ** 2).sum(axis=1) / nz
            pl["thu_xz"] = (ft * f1).sum(axis=1) / nz
            pl["thw_xz"] = pts["thw_zstar01"].sum(axis=1) / nz
            pl["chi_xz"] = pts["chi_zstar01"].sum(axis=1) / nz
        out["_planes"] = {k: v[1:ny + 1] for k, v in pl.items()}
        pts["u1"], pts["u2"], pts["u3"], pts["th"] = u1, u2, u3, th
    return out


def movie(o, nxm, nym, nzm, time=0.0):
    """The movie output of save_stats_chan(movie = .true.) for the oracle's
    state: (planes, slices) keyed by the reference's group names, in the files'
    C layouts (mean_xz planes [Nyp][Nx]; slices <f>_xz [Nyp][Nx] at z = nzm,
    <f>_xy [Nz][Nx] at row nym (absent for nym = -1), <f>_yz [Nyp][Nz] at
    x = nxm), for/diagnostics.f90:408-441, 485-617, 1041-1070, 1566-1727."""
    cfg = o.p.cfg
    ny = cfg.ny
    out = save_stats(o, time, points=True)
    pts = out["_points"]
    u2 = pts["u2"]
    w = np.zeros_like(u2)
    w[:ny + 1] = 0.5 * (u2[:ny + 1] + u2[1:ny + 2])
    nut = o.download_les()[0] if cfg.use_les else np.zeros_like(u2)
    fields = {"u": pts["u1"], "w": w, "v": pts["u3"], "nu_t": nut, "epsilon": pts["epsilon_zstar"],
              "omegaX": pts["omega_x"], "omegaZ": pts["omega_y"], "omegaY": pts["omega_z"]}
    for n, t in enumerate(pts["th"]):
        fields[f"th{n + 1}"] = t
    sl = {}
    J = slice(1, ny + 1)
    for name, a in fields.items():
        sl[name + "_xz"] = a[J, nzm, :]
        if nym >= 1:
            sl[name + "_xy"] = a[nym, :, :]
        sl[name + "_yz"] = a[J, :, nxm]
    return out["_planes"], sl


# the reference's pi: 4 atan(1.0) in single precision (for/fft.f90:147)
PI_F32 = float(np.float32(4.0) * np.arctan(np.float32(1.0)))
ZSTAR_FIELDS = ["shear2_zstar", "epsilon_zstar", "TKE_zstar", "KE_zstar"]


def zstar_bins(cfg, dyf, th0, fields, hx):
    """compute_BPE and Bin_Ystar_and_Write (for/diagnostics.f90:1733-1852,
    1859-1953) on given point fields: th0 and each field are [j][z][x] arrays
    over rows 0..Nyp+1.  Returns {'th1PDF', 'th1bin', 'BPE', name: binned mean}."""
    nx = th0.shape[2]
    nb = nx
    jlo, jhi = cfg.jstart_th[0], cfg.jend_th[0]
    gxv = (np.arange(nx) * cfg.lx) / nx                       # gx(i) = (i Lx) / Nx
    s1 = th0[jlo:jhi + 1] + cfg.dthdx[0] * (gxv - 0.5 * cfg.lx)[None, None, :]
    if hx:
        thmin = -(cfg.dthdx[0] * 0.5 * cfg.lx * float(np.float32(1.1)))  # the literal 1.1 is single
        thmax = cfg.dthdx[0] * 0.5 * cfg.lx * float(np.float32(1.1))
        dth = (thmax - thmin) / nb + 1.0e-14
        th_bin = thmin + np.arange(nb + 1) * dth
        b = np.minimum(np.trunc((s1 - thmin) / dth).astype(np.int64), nb - 1)
        b = np.maximum(b, 0)  # below thmin the reference indexes outside its array; here the first bin
    else:
        den = cfg.ro_inv + 1.0e-14
        th_bin = -den * np.cos(np.arange(nb + 1) * PI_F32 / nb)
        c = np.minimum(np.maximum(s1, -cfg.ro_inv), cfg.ro_inv)
        b = np.minimum(np.trunc(np.arccos(-c / den) * nb / PI_F32).astype(np.int64), nb - 1)
    w = np.broadcast_to(dyf[jlo:jhi + 1][:, None, None], s1.shape)
    pdf_raw = np.bincount(b.ravel(), weights=w.ravel(), minlength=nb)
    res = {}
    with np.errstate(invalid="ignore", divide="ignore"):
        for name, f in fields.items():
            sums = np.bincount(b.ravel(), weights=(f[jlo:jhi + 1] * w).ravel(), minlength=nb)
            res[name] = sums / pdf_raw
    pdf = pdf_raw / pdf_raw.sum()
    yr = np.zeros(nb + 1)
    for i in range(1, nb + 1):
        yr[i] = yr[i - 1] + pdf[i - 1]
    yr = yr * cfg.ly
    bpe = 0.0
    for i in range(nb):
        bpe = bpe - (0.5 * (th_bin[i + 1] + th_bin[i]) * 0.5 * (yr[i + 1] + yr[i])) * (yr[i + 1] - yr[i]) / cfg.ly
    res.update({"th1PDF": pdf, "th1bin": th_bin, "BPE": bpe, "_bin": b})
    return res


def zstar(o, time=0.0):
    """The z* statistics of the oracle's state (after the BC posts of save_stats)."""
    cfg = o.p.cfg
    out = save_stats(o, time, points=True)
    pts = out["_points"]
    names = ZSTAR_FIELDS + [f"{b}{n + 1:02d}" for n in range(cfg.n_th) for b in ("thw_zstar", "chi_zstar")]
    return zstar_bins(cfg, o.p.met["dyf"], pts["th0"], {k: pts[k] for k in names}, homogeneous_x(cfg))
