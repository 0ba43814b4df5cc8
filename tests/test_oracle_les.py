"""Pins for the oracle's LES branch: the AMD closure of les_chan / les_chan_th,
les_model_type 5 (for/les.f90:551-901, 1201-1395).

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
