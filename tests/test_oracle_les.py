"""Pins for the oracle's LES branch: les_chan / les_chan_th for les_model_type 5
(AMD of the scaled gradient, for/les.f90:551-901, 1201-1395), 1 (constant
Smagorinsky, 77-254) and 4 (AMD weighted by the filter widths, 260-550).

The reference ships no LES test data, so each check compares the oracle against
something of a different form:
  * closed forms: the AMD coefficient of a pure shear is zero, and for a field whose
    resolved gradient is diagonal it is max(0, -(g1^3 + g3^3)) / (g1^2 + g3^2);
    the scalar coefficient of theta = theta(x) is max(0, -du1/dx);
  * the tensor form of the model, -tr(G G^T S) / |G|^2 with G the filter-scaled
    gradient and S its symmetric part (the oracle expands it into 30 products);
  * the SGS force, rebuilt with numpy's FFT from the closed-form stress;
  * the spin-up gate of rk_chan_1 (for/solver.f90:615).
"""
import math

import numpy as np
import pytest

import oracle
from diablo3_amd import configs, layout
from diablo3_amd.ic import FourierState
from diablo3_amd.problem import Problem

C_AMD = 0.2887
BETA_SGS = 3.0


def les_cfg(nx=16, ny=17, nz=16):
    cfg = configs.small_dns(nx, ny, nz)
    cfg.use_les = True
    cfg.les_model_type = 5
    return cfg


def delta_fac(cfg, h):
    """-0.5 * delta(j): the factor between the AMD coefficient and nu_t."""
    dx, dz = cfg.lx / cfg.nx * BETA_SGS, cfg.lz / cfg.nz * BETA_SGS
    return 3.0 * C_AMD ** 2 / (1 / dx ** 2 + 1 / (2 * h) ** 2 + 1 / dz ** 2)


def state_from_physical(p, u1, u2, u3, th):
    cfg = p.cfg
    f = lambda a: layout.fft_xz_to_fourier(cfg, a)
    return FourierState(f(u1), f(u2), f(u3), layout.compact_zeros(cfg), [f(th)])


def grids(cfg):
    x = np.arange(cfg.nx) * cfg.lx / cfg.nx
    z = np.arange(cfg.nz) * cfg.lz / cfg.nz
    return x, z


def run_substep(p, s, time_step=1, create_new_flow=0):
    o = oracle.Oracle(p, threads=2)
    o.upload_state(s)
    o.rk_substep(1, time=0.0, time_step=time_step, create_new_flow=create_new_flow)
    return o


def test_pure_shear_has_no_eddy_viscosity():
    cfg = les_cfg()
    p = Problem(cfg)
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    y = p.met["gyf"][:, None, None]
    u1 = np.broadcast_to(np.sin(1.3 * y) + 0.2 * y ** 2, shp).copy()
    th = np.broadcast_to(0.7 * y + 0.1 * np.cos(2 * y), shp).copy()
    o = run_substep(p, state_from_physical(p, u1, np.zeros(shp), np.zeros(shp), th))
    nt, kt = o.download_les()
    assert np.all(nt == 0.0)
    assert np.all(kt[0] == 0.0)


def test_diagonal_strain_closed_form():
    cfg = les_cfg(32, 17, 16)
    p = Problem(cfg)
    x, z = grids(cfg)
    a, b, c = 0.8, -0.5, 0.3
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    u1 = np.broadcast_to(a * np.sin(2 * math.pi * x / cfg.lx)[None, None, :], shp).copy()
    u3 = np.broadcast_to(b * np.sin(2 * math.pi * z / cfg.lz)[None, :, None], shp).copy()
    th = np.broadcast_to(c * np.sin(2 * math.pi * x / cfg.lx)[None, None, :], shp).copy()
    o = run_substep(p, state_from_physical(p, u1, np.zeros(shp), u3, th))
    nt, kt = o.download_les()
    # spectral derivatives with the float32-pi wavenumbers of init_fft
    g1 = (a * p.kx[1] * np.cos(2 * math.pi * x / cfg.lx))[None, :]
    g3 = (b * p.kz[1] * np.cos(2 * math.pi * z / cfg.lz))[:, None]
    num = -(g1 ** 3 + g3 ** 3)
    coef = np.where(num > 0, np.maximum(num, 0) / (g1 ** 2 + g3 ** 2), 0.0)
    coef_th = np.where(-g1 > 0, -g1, 0.0) + 0 * g3
    dy = p.met["dy"]
    for j in range(3, cfg.ny - 1):
        np.testing.assert_allclose(nt[j], delta_fac(cfg, dy[j]) * coef, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(kt[0][j], delta_fac(cfg, dy[j]) * coef_th, rtol=1e-12, atol=1e-15)
    # walls: ghost_les_mpi zeroes j = 1, 2, Nyp, Nyp+1
    for j in (0, 1, 2, cfg.ny, cfg.ny + 1):
        assert np.all(nt[j] == 0) and np.all(kt[0][j] == 0)
    assert nt.max() > 1e-4  # the closed form is not trivially zero here


def test_sgs_force_from_closed_form_stress():
    """cf(LES) - cf(no LES) on interior rows equals the divergence of the
    closed-form stress, transformed with numpy's FFT (a different transform)."""
    cfg = les_cfg(32, 17, 16)
    p = Problem(cfg)
    x, z = grids(cfg)
    a, b = 0.8, -0.5
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    u1 = np.broadcast_to(a * np.sin(2 * math.pi * x / cfg.lx)[None, None, :], shp).copy()
    u3 = np.broadcast_to(b * np.sin(2 * math.pi * z / cfg.lz)[None, :, None], shp).copy()
    zero = np.zeros(shp)
    s = state_from_physical(p, u1, zero, u3, zero)
    o_les = run_substep(p, s)
    cfg0 = configs.small_dns(32, 17, 16)
    o_dns = run_substep(Problem(cfg0), s)
    (cf1, cf2, cf3), _ = o_les.download_rk()
    (df1, df2, df3), _ = o_dns.download_rk()
    g1 = (a * p.kx[1] * np.cos(2 * math.pi * x / cfg.lx))[None, :]
    g3 = (b * p.kz[1] * np.cos(2 * math.pi * z / cfg.lz))[:, None]
    num = -(g1 ** 3 + g3 ** 3)
    coef = np.where(num > 0, num / (g1 ** 2 + g3 ** 2), 0.0)
    kx = p.kx[:cfg.kxr][None, :]
    kz = p.kz[:, None]
    for j in range(3, cfg.ny - 1):
        d = -2.0 * delta_fac(cfg, p.met["dyf"][j])  # delta_yf(j)
        t11 = layout.fft_xz_to_fourier(cfg, (d * coef * g1)[None])[0]
        t33 = layout.fft_xz_to_fourier(cfg, (d * coef * g3)[None])[0]
        np.testing.assert_allclose(cf1[j] - df1[j], -1j * kx * t11, rtol=0, atol=1e-12)
        np.testing.assert_allclose(cf3[j] - df3[j], -1j * kz * t33, rtol=0, atol=1e-12)
        np.testing.assert_allclose(cf2[j] - df2[j], 0.0, atol=1e-13)
    assert np.abs(cf1[5] - df1[5]).max() > 1e-6


def _tensor_amd(G, S):
    """-tr(G G^T S) / |G|^2, clipped at zero; G, S are [..., 3, 3]."""
    num = -np.einsum("...ij,...kj,...ki->...", G, G, S)
    den = np.einsum("...ij,...ij->...", G, G)
    return np.where(num > 0, num / np.where(den > 0, den, 1.0), 0.0)


def test_random_field_matches_tensor_form():
    cfg = les_cfg(16, 17, 16)
    p = Problem(cfg)
    s = p.initial_state(seed=3, theta_noise=1e-2)
    rng = np.random.default_rng(5)
    # add a large random smooth field so that the coefficient is far from zero
    for a in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        a[2:cfg.ny - 1, 1:4, 1:4] += 0.05 * (rng.standard_normal((cfg.ny - 3, 3, 3)) +
                                             1j * rng.standard_normal((cfg.ny - 3, 3, 3)))
    o = run_substep(p, s)
    nt, kt = o.download_les()
    # gradients of the uploaded state (planes 3..Ny-1 are not touched by the BC posts)
    phys = lambda c: layout.fft_xz_to_physical(cfg, c)
    ikx = 1j * p.kx[:cfg.kxr][None, None, :]
    ikz = 1j * p.kz[None, :, None]
    dy, dyf = p.met["dy"], p.met["dyf"]
    u = [s.cu1, s.cu2, s.cu3]
    dx_ = [phys(ikx * c) for c in u]   # d/dx of u1, u2, u3 (at each field's own points)
    dz_ = [phys(ikz * c) for c in u]
    up = [phys(c) for c in u]
    th = phys(s.cth[0])
    dthx, dthz = phys(ikx * s.cth[0]), phys(ikz * s.cth[0])
    Dx, Dz = cfg.lx / cfg.nx * BETA_SGS, cfg.lz / cfg.nz * BETA_SGS
    checked = 0
    for j in range(3, cfg.ny - 1):
        Dy = 2 * dy[j]
        w = lambda f: (f[j] * dyf[j - 1] + f[j - 1] * dyf[j]) / (2 * dy[j])  # GYF -> GY
        du2dy = lambda jj: (up[1][jj + 1] - up[1][jj]) / dyf[jj]
        grad = np.zeros((cfg.nz, cfg.nx, 3, 3))  # grad[..., i, k] = d u_i / d x_k at GY j
        grad[..., 0, 0] = w(dx_[0])
        grad[..., 0, 1] = (up[0][j] - up[0][j - 1]) / dy[j]
        grad[..., 0, 2] = w(dz_[0])
        grad[..., 1, 0] = dx_[1][j]
        grad[..., 1, 1] = (du2dy(j) * dyf[j - 1] + du2dy(j - 1) * dyf[j]) / (2 * dy[j])
        grad[..., 1, 2] = dz_[1][j]
        grad[..., 2, 0] = w(dx_[2])
        grad[..., 2, 1] = (up[2][j] - up[2][j - 1]) / dy[j]
        grad[..., 2, 2] = w(dz_[2])
        D = np.array([Dx, Dy, Dz])
        G = grad * D[None, None, None, :] / D[None, None, :, None]   # G_ik = du_i/dx_k D_k / D_i
        S = 0.5 * (G + np.swapaxes(G, -1, -2))
        np.testing.assert_allclose(nt[j], delta_fac(cfg, dy[j]) * _tensor_amd(G, S), rtol=1e-11, atol=1e-17)
        # scalar: -sum_ki D_k^2 du_i/dx_k dth/dx_k dth/dx_i / sum_k (D_k dth/dx_k)^2
        gt = np.stack([w(dthx), (th[j] - th[j - 1]) / dy[j], w(dthz)], axis=-1)
        num = -np.einsum("k,...ik,...k,...i->...", D ** 2, grad, gt, gt)
        den = np.einsum("k,...k->...", D ** 2, gt ** 2)
        ref = np.where(num > 0, num / den, 0.0)
        np.testing.assert_allclose(kt[0][j], delta_fac(cfg, dy[j]) * ref, rtol=1e-11, atol=1e-17)
        checked += int((nt[j] > 0).sum() + (kt[0][j] > 0).sum())
    assert checked > 100  # both clipped and active points are exercised


def test_spin_up_gate():
    """With create_new_flow the SGS model stays off for time_step <= 100."""
    cfg = les_cfg(16, 17, 16)
    p = Problem(cfg)
    s = p.initial_state(seed=1, theta_noise=1e-3)
    o_dns = run_substep(Problem(configs.small_dns(16, 17, 16)), s)
    for ts, same in ((100, True), (101, False)):
        o = run_substep(p, s, time_step=ts, create_new_flow=1)
        a, b = o.download_state(), o_dns.download_state()
        eq = all(np.array_equal(x, y) for x, y in ((a.cu1, b.cu1), (a.cu2, b.cu2), (a.cu3, b.cu3),
                                                   (a.cth[0], b.cth[0])))
        assert eq == same, ts


# ---------------------------------------------------------------- types 1 and 4
C_SMAG = 0.17


def les_cfg_model(model, nx=16, ny=17, nz=16):
    cfg = les_cfg(nx, ny, nz)
    cfg.les_model_type = model
    return cfg


def test_type1_smagorinsky_closed_forms():
    """Constant Smagorinsky (for/les.f90:77-254): nu_t = c^2 (Dx Dy Dz)^(2/3) |S| at GY,
    |S| = |U'| for a pure shear U(y), sqrt(2 (g1^2 + g3^2)) for a diagonal gradient;
    kappa_t = nu_t (for/les.f90:1118-1126)."""
    cfg = les_cfg_model(1, 32, 17, 16)
    p = Problem(cfg)
    dy, gyf = p.met["dy"], p.met["gyf"]
    bx, bz = cfg.lx / cfg.nx * BETA_SGS, cfg.lz / cfg.nz * BETA_SGS
    fac = lambda j: C_SMAG ** 2 * (bx * dy[j] * 2.0 * bz) ** (2.0 / 3.0)
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    y = gyf[:, None, None]
    u1 = np.broadcast_to(0.3 * y ** 2 + 0.1 * y, shp).copy()
    zero = np.zeros(shp)
    o = run_substep(p, state_from_physical(p, u1, zero, zero, zero))
    nt, kt = o.download_les()
    for j in range(3, cfg.ny - 1):
        dudy = (u1[j, 0, 0] - u1[j - 1, 0, 0]) / dy[j]
        np.testing.assert_allclose(nt[j], fac(j) * abs(dudy), rtol=1e-12)
        np.testing.assert_array_equal(kt[0][j], nt[j])
    x, z = grids(cfg)
    a, b = 0.8, -0.5
    u1 = np.broadcast_to(a * np.sin(2 * math.pi * x / cfg.lx)[None, None, :], shp).copy()
    u3 = np.broadcast_to(b * np.sin(2 * math.pi * z / cfg.lz)[None, :, None], shp).copy()
    o = run_substep(p, state_from_physical(p, u1, zero, u3, zero))
    nt, _ = o.download_les()
    g1 = (a * p.kx[1] * np.cos(2 * math.pi * x / cfg.lx))[None, :]
    g3 = (b * p.kz[1] * np.cos(2 * math.pi * z / cfg.lz))[:, None]
    for j in range(3, cfg.ny - 1):
        np.testing.assert_allclose(nt[j], fac(j) * np.sqrt(2 * (g1 ** 2 + g3 ** 2)), rtol=1e-12, atol=1e-16)


def test_type4_shear_and_diagonal_closed_forms():
    cfg = les_cfg_model(4, 32, 17, 16)
    p = Problem(cfg)
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    y = p.met["gyf"][:, None, None]
    zero = np.zeros(shp)
    u1 = np.broadcast_to(np.sin(1.3 * y), shp).copy()
    nt, kt = run_substep(p, state_from_physical(p, u1, zero, zero, zero)).download_les()
    assert np.all(nt == 0.0) and np.all(kt[0] == 0.0)
    x, z = grids(cfg)
    a, b = 0.8, -0.5
    u1 = np.broadcast_to(a * np.sin(2 * math.pi * x / cfg.lx)[None, None, :], shp).copy()
    u3 = np.broadcast_to(b * np.sin(2 * math.pi * z / cfg.lz)[None, :, None], shp).copy()
    nt, kt = run_substep(p, state_from_physical(p, u1, zero, u3, zero)).download_les()
    dX = (cfg.lx / cfg.nx * BETA_SGS) ** 2
    dZ = (cfg.lz / cfg.nz * BETA_SGS) ** 2
    g1 = (a * p.kx[1] * np.cos(2 * math.pi * x / cfg.lx))[None, :]
    g3 = (b * p.kz[1] * np.cos(2 * math.pi * z / cfg.lz))[:, None]
    num = -(dX * g1 ** 3 + dZ * g3 ** 3)
    ref = C_AMD ** 2 * np.where(num > 0, num / (g1 ** 2 + g3 ** 2), 0.0)
    for j in range(3, cfg.ny - 1):
        np.testing.assert_allclose(nt[j], ref, rtol=1e-12, atol=1e-16)
        np.testing.assert_array_equal(kt[0][j], nt[j])
    assert nt.max() > 0


def test_type4_random_field_matches_tensor_form():
    """Type 4 is an anisotropic AMD of the unscaled gradient G = grad u:
    -(1/2) sum_k D_k^2 (G_ik G_jk + G_ki G_kj) S_ij / |G|^2, D = (dx beta, 2 dy(j), dz beta)."""
    cfg = les_cfg_model(4, 16, 17, 16)
    p = Problem(cfg)
    s = p.initial_state(seed=3, theta_noise=1e-2)
    rng = np.random.default_rng(7)
    for arr in (s.cu1, s.cu2, s.cu3):
        arr[2:cfg.ny - 1, 1:4, 1:4] += 0.05 * (rng.standard_normal((cfg.ny - 3, 3, 3)) +
                                               1j * rng.standard_normal((cfg.ny - 3, 3, 3)))
    o = run_substep(p, s)
    nt, _ = o.download_les()
    phys = lambda c: layout.fft_xz_to_physical(cfg, c)
    ikx = 1j * p.kx[:cfg.kxr][None, None, :]
    ikz = 1j * p.kz[None, :, None]
    dy, dyf = p.met["dy"], p.met["dyf"]
    u = [s.cu1, s.cu2, s.cu3]
    dx_ = [phys(ikx * c) for c in u]
    dz_ = [phys(ikz * c) for c in u]
    up = [phys(c) for c in u]
    active = 0
    for j in range(3, cfg.ny - 1):
        w = lambda f: (f[j] * dyf[j - 1] + f[j - 1] * dyf[j]) / (2 * dy[j])
        du2dy = lambda jj: (up[1][jj + 1] - up[1][jj]) / dyf[jj]
        G = np.zeros((cfg.nz, cfg.nx, 3, 3))
        G[..., 0, 0] = w(dx_[0]); G[..., 0, 1] = (up[0][j] - up[0][j - 1]) / dy[j]; G[..., 0, 2] = w(dz_[0])
        G[..., 1, 0] = dx_[1][j]
        G[..., 1, 1] = (du2dy(j) * dyf[j - 1] + du2dy(j - 1) * dyf[j]) / (2 * dy[j])
        G[..., 1, 2] = dz_[1][j]
        G[..., 2, 0] = w(dx_[2]); G[..., 2, 1] = (up[2][j] - up[2][j - 1]) / dy[j]; G[..., 2, 2] = w(dz_[2])
        D2 = np.array([(cfg.lx / cfg.nx * BETA_SGS) ** 2, (2 * dy[j]) ** 2, (cfg.lz / cfg.nz * BETA_SGS) ** 2])
        S = 0.5 * (G + np.swapaxes(G, -1, -2))
        # the expansion of for/les.f90:276-343 is the mean of the two contractions
        num = -0.5 * (np.einsum("k,...ik,...jk,...ij->...", D2, G, G, S) +
                      np.einsum("k,...ki,...kj,...ij->...", D2, G, G, S))
        den = np.einsum("...ij,...ij->...", G, G)
        ref = C_AMD ** 2 * np.where(num > 0, num / den, 0.0)
        np.testing.assert_allclose(nt[j], ref, rtol=1e-11, atol=1e-18)
        active += int((ref > 0).sum())
    assert active > 100
