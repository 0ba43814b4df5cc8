"""Pins for the oracle's restatement of the energy-budget statistics
(oracle/diagnostics.py, save_stats_chan of for/diagnostics.f90).

The reference holds no stats data, so the physical-space plane sums are checked
against Parseval's theorem on the Fourier state, a different formulation: a plane
(or z-line) average removes exactly the kx = kz = 0 mode (or the kz = 0 row), so
every fluctuation second moment is a weighted sum of |c|^2 over the remaining
modes (weight 2 for kx > 0, the Hermitian half the state omits).
"""
import numpy as np
import pytest

import oracle
from oracle import diagnostics
from diablo3_amd import configs
from diablo3_amd.problem import Problem


def state(cfg_fn, seed=4):
    cfg = cfg_fn()
    p = Problem(cfg)
    o = oracle.Oracle(p, threads=4)
    s = p.initial_state(seed=seed, theta_noise=1e-2)
    rng = np.random.default_rng(seed)
    sh = (cfg.ny + 2, cfg.kzr, cfg.kxr)
    for a in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        a += 1e-2 * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
        a[:, :, 0] = a[:, :, 0].real + 0j * a[:, :, 0]  # keep kx = 0 columns Hermitian-consistent below
        # Hermitian symmetry of the kx = 0 column in kz (a real field)
        nkz = cfg.nkz
        a[:, nkz + 1:, 0] = np.conj(a[:, nkz:0:-1, 0])
        a[:, 0, 0] = a[:, 0, 0].real
    o.upload_state(s)
    o.rem_div()
    return p, o


def weights(cfg):
    w = np.full(cfg.kxr, 2.0)
    w[0] = 1.0
    return w[None, :]


def parseval(cfg, c, krow_from):
    """sum over k >= krow_from (or all but (0,0) when krow_from == 0) of w |c|^2, per plane."""
    w = weights(cfg)
    a = np.abs(c) ** 2 * w[None]
    if krow_from == 0:
        a = a.copy()
        a[:, 0, 0] = 0.0
        return a.sum(axis=(1, 2))
    return a[:, krow_from:].sum(axis=(1, 2))


@pytest.mark.parametrize("cfg_fn,hx", [(lambda: configs.small_dns(16, 17, 16), True),
                                        (lambda: configs.front_les(16, 17, 16, 5000.0, 2e-4, 1e-2), False)])
def test_second_moments_parseval(cfg_fn, hx):
    p, o = state(cfg_fn)
    cfg = p.cfg
    assert diagnostics.homogeneous_x(cfg) == hx
    st = diagnostics.save_stats(o)
    s = o.download_state()  # after the BC posts save_stats applied
    k0 = 0 if hx else 1
    J = slice(1, cfg.ny + 1)
    for name, c in (("urms", s.cu1), ("wrms", s.cu3), ("vrms", s.cu2)):
        np.testing.assert_allclose(st[name][J] ** 2, parseval(cfg, c, k0)[J], rtol=1e-11)
    w = weights(cfg)
    cross = (s.cu1 * np.conj(s.cu3)).real * w[None]
    ref = cross[:, k0:].sum(axis=(1, 2)) if k0 else cross.sum(axis=(1, 2)) - cross[:, 0, 0]
    np.testing.assert_allclose(st["uw"][J], ref[J], rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(st["thrms01"][J] ** 2, parseval(cfg, s.cth[0], k0)[J], rtol=1e-11)


def test_dissipation_and_gradients_parseval():
    """epsilon and the scalar gradient variance of an x-homogeneous state, as mode
    sums: x and z derivatives weight |c|^2 by kx^2, kz^2; y differences act per mode."""
    p, o = state(lambda: configs.small_dns(16, 17, 16))
    cfg, met = p.cfg, p.met
    st = diagnostics.save_stats(o)
    s = o.download_state()
    dy, dyf, gy, gyf = met["dy"], met["dyf"], met["gy"], met["gyf"]
    kx2 = (p.kx[:cfg.kxr] ** 2)[None, None, :]
    kz2 = (p.kz ** 2)[None, :, None]
    w = weights(cfg)[None]
    def ms(c):  # plane mean square of the fluctuation of spectral field c
        a = np.abs(c) ** 2 * w
        a[:, 0, 0] = 0.0
        return a.sum(axis=(1, 2))
    ny = cfg.ny
    eps = np.zeros(ny + 2)
    for j in range(2, ny + 1):
        it = lambda v: (dyf[j - 1] * v[j] + dyf[j] * v[j - 1]) / (2 * dy[j])
        gx1 = (np.abs(s.cu1) ** 2 * kx2 * w).sum(axis=(1, 2))
        gx3 = (np.abs(s.cu3) ** 2 * kx2 * w).sum(axis=(1, 2))
        gz1 = (np.abs(s.cu1) ** 2 * kz2 * w).sum(axis=(1, 2))
        gz3 = (np.abs(s.cu3) ** 2 * kz2 * w).sum(axis=(1, 2))
        gx2 = (np.abs(s.cu2) ** 2 * kx2 * w).sum(axis=(1, 2))
        gz2 = (np.abs(s.cu2) ** 2 * kz2 * w).sum(axis=(1, 2))
        e = it(gx1) + gx2[j] + it(gx3) + it(gz1) + gz2[j] + it(gz3)
        e += ms((s.cu1[j:j + 1] - s.cu1[j - 1:j]) / dy[j])[0]
        e += ms((s.cu2[j + 1:j + 2] - s.cu2[j - 1:j]) / (gy[j + 1] - gy[j - 1]))[0]
        e += ms((s.cu3[j:j + 1] - s.cu3[j - 1:j]) / dy[j])[0]
        eps[j] = cfg.nu * e
    np.testing.assert_allclose(st["epsilon"][2:ny + 1], eps[2:ny + 1], rtol=1e-10)
    # (grad th + background)^2: mode sums plus the squared background gradients
    ct = s.cth[0]
    gxt = (np.abs(ct) ** 2 * kx2 * w).sum(axis=(1, 2))
    gzt = (np.abs(ct) ** 2 * kz2 * w).sum(axis=(1, 2))
    for j in range(1, ny + 1):
        d = (ct[j + 1] - ct[j - 1]) / (gyf[j + 1] - gyf[j - 1])
        gyt = (np.abs(d) ** 2 * w[0]).sum()
        ref = gxt[j] + gyt + gzt[j] + cfg.dthdx[0] ** 2 + cfg.dthdz[0] ** 2
        np.testing.assert_allclose(st["pe_diss01"][j], ref, rtol=1e-10)
    # the mean profiles are the (0,0) modes
    np.testing.assert_array_equal(st["ume"], s.cu1[:, 0, 0].real)
    np.testing.assert_array_equal(st["thme01"], s.cth[0][:, 0, 0].real)


@pytest.mark.parametrize("hx", [True, False])
def test_zstar_bins_closed_forms(hx):
    """The oracle's z* binning (compute_BPE, Bin_Ystar_and_Write) on synthetic
    fields: each point lies between the edges of its bin; the PDF sums to 1; a
    constant field bins to that constant in every occupied bin (NaN in the empty
    ones); with s1 constant all weight is in one bin b and
    BPE = -Ly (edge_b + edge_b+1) / 4."""
    from oracle import diagnostics as dg
    cfg = configs.front_les(16, 9, 8, 5000.0, 2e-4, 1e-2)
    p = Problem(cfg)
    dyf = p.met["dyf"]
    ny2, nz, nx = cfg.ny2, cfg.nz, cfg.nx
    rng = np.random.default_rng(4)
    gx = (np.arange(nx) * cfg.lx) / nx
    # random scalar spanning the bins, plus two constant fields
    # (x homogeneous: inside the uniform bins, |s1| < 0.55 dTHdX Lx; else past the
    # ends too, where the acos bins clip)
    th0 = rng.uniform(-1.0, 1.0, (ny2, nz, nx)) * (0.6 * cfg.ro_inv if not hx else 0.04 * cfg.dthdx[0] * cfg.lx)
    r = dg.zstar_bins(cfg, dyf, th0, {"c": np.full((ny2, nz, nx), 2.5)}, hx)
    jlo, jhi = cfg.jstart_th[0], cfg.jend_th[0]
    s1 = th0[jlo:jhi + 1] + cfg.dthdx[0] * (gx - 0.5 * cfg.lx)
    b, e = r["_bin"], r["th1bin"]
    inside = (s1 >= e[b] - 1e-12) & ((s1 <= e[b + 1] + 1e-12) | (b == nx - 1))
    if not hx:
        inside |= (s1 <= -cfg.ro_inv) | (s1 >= cfg.ro_inv)  # clipped into the end bins
    assert inside.all()
    assert abs(r["th1PDF"].sum() - 1.0) < 1e-14
    occ = r["th1PDF"] > 0
    assert np.allclose(r["c"][occ], 2.5, rtol=1e-13, atol=0) and np.all(np.isnan(r["c"][~occ]))
    # s1 constant: all the weight in one bin
    th1 = np.broadcast_to(0.1 * cfg.ro_inv - cfg.dthdx[0] * (gx - 0.5 * cfg.lx), (ny2, nz, nx)).copy()
    r1 = dg.zstar_bins(cfg, dyf, th1, {}, hx)
    k = int(np.argmax(r1["th1PDF"]))
    assert r1["th1PDF"][k] == 1.0 and np.count_nonzero(r1["th1PDF"]) == 1
    e1 = r1["th1bin"]
    assert abs(r1["BPE"] - (-cfg.ly * (e1[k] + e1[k + 1]) / 4.0)) <= 1e-14 * cfg.ly * abs(e1).max()


def test_vorticity_closed_form():
    """compute_Vorticity in the oracle on a pure shear u1 = a gyf (mean mode):
    omega_z = a (gyf(j+1) - gyf(j-1)) / (2 dyf(j)) on every point, omega_x the
    thermal-wind shear alone, omega_y = 0 (rows clear of the BC posts)."""
    cfg = configs.front_les(16, 17, 8, 5000.0, 2e-4, 1e-2)
    cfg.use_les = False
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s = p.initial_state(seed=1)
    for arr in (s.cu1, s.cu2, s.cu3, s.cp, *s.cth):
        arr[:] = 0
    gyf, dyf = p.met["gyf"], p.met["dyf"]
    a = 0.75
    s.cu1[:, 0, 0] = a * gyf
    o.upload_state(s)
    out = diagnostics.save_stats(o)
    twf = (1.0 / (cfg.ro_inv / cfg.delta)) * cfg.dthdx[0]
    J = slice(3, cfg.ny - 1)
    want_z = np.abs(a * (gyf[4:cfg.ny] - gyf[2:cfg.ny - 2]) / (2.0 * dyf[J]))
    assert np.allclose(out["omega_z"][J], want_z, rtol=1e-13, atol=0)
    assert np.allclose(out["omega_x"][J], abs(twf), rtol=1e-13, atol=0)
    assert np.abs(out["omega_y"][J]).max() < 1e-14


def test_ftx_uu_closed_form():
    """FTx_uu in the oracle: a u1 with kz = 0 modes of constant amplitude c_i in y
    (row 1 then zeroed by the bottom no-slip BC post of save_stats; the free-slip
    top keeps row Nyp) gives |c_i|^2 (dy(2)/2 + dy(3) + ... + dy(Nyp)) / Ly per
    mode (the trapezoid weights of a constant), zero past Nkx."""
    cfg = configs.small_dns(16, 17, 8)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s = p.initial_state(seed=1)
    for arr in (s.cu1, s.cu2, s.cu3, s.cp, *s.cth):
        arr[:] = 0
    c = np.array([0.3, 0.2 + 0.1j, -0.05j, 0.07, 0.0, 0.01 + 0.02j])[:cfg.kxr]
    s.cu1[:, 0, :len(c)] = c[None, :]
    o.upload_state(s)
    out = diagnostics.save_stats(o)
    dy = p.met["dy"]
    after = o.download_state().cu1
    assert np.all(after[1, 0] == 0) and np.all(after[2:cfg.ny + 1, 0, :len(c)] == c[None, :])
    want = np.zeros(cfg.nx // 2)
    want[:len(c)] = np.abs(c) ** 2 * (0.5 * dy[2] + dy[3:cfg.ny + 1].sum()) / cfg.ly
    assert np.allclose(out["FTx_uu"], want, rtol=1e-13, atol=1e-18)
