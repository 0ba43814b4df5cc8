"""Pins for the CPU oracle.

The reference ships no tests, golden vectors or fixtures for this path, so every
oracle function is checked here against something independent of it: numpy's FFT
and a brute-force DFT, dense linear solves, closed forms (float32-pi wavenumbers,
the RKW3 stability polynomial 1 + z + z^2/2 + z^3/6), and invariants
(divergence at round-off after projection, translation equivariance of a step).
"""
import math

import numpy as np
import pytest

import oracle
from diablo3_amd import configs, grid, layout
from diablo3_amd.config import Config, DIRICHLET, NEUMANN
from diablo3_amd.ic import FourierState
from diablo3_amd.problem import Problem

PI_F32 = 3.1415927410125732  # float(np.float32(math.pi)), the value 4.*atan(1.0) takes in real(4)


# ---------------------------------------------------------------- wavenumbers, grid
def test_wavenumbers_closed_form():
    nx, nz, lx, lz = 64, 48, 2 * math.pi, 8 * math.pi
    kx, kz = oracle.wavenumbers(nx, nz, lx, lz)
    hx, hz = grid.wavenumbers(nx, nz, lx, lz)
    ex = np.array([i * (2 * PI_F32) / lx for i in range(nx // 2)])
    nkz = nz // 3
    ez = np.array([k * (2 * PI_F32) / lz for k in range(nkz + 1)] +
                  [-(nkz + 1 - m) * (2 * PI_F32) / lz for m in range(1, nkz + 1)])
    assert np.array_equal(kx, ex) and np.array_equal(hx, ex)
    assert np.array_equal(kz, ez) and np.array_equal(hz, ez)
    # the float32 pi makes kx(1) Lx / 2pi = 1 + 2.8e-8, not 1
    assert abs(kx[1] * lx / (2 * math.pi) - 1 - 2.78e-8) < 1e-9


def test_grid_walls_and_metrics():
    for n, cs in ((65, 1.75), (129, 1.75), (17, 1.5)):
        g = grid.create_grid(n, 2.0, cs, 1)
        assert len(g) == n + 1 and np.all(np.diff(g) > 0)
        m = grid.metrics(g)
        mo = oracle.grid_metrics(g)
        for k in m:
            assert np.array_equal(m[k], mo[k]), k
        # the script puts GYF(1) and GYF(N) on the walls y = -1, +1
        assert abs(m["gyf"][1] + 1.0) < 1e-14 and abs(m["gyf"][n] - 1.0) < 1e-14
        # symmetric tanh stretching
        assert np.allclose(g, -g[::-1], atol=1e-14)
        assert np.all(m["dy"][1:] > 0) and np.all(m["dyf"] > 0)
        # gyf is the midpoint of gy, dyf the gy spacing
        assert np.allclose(m["gyf"][1:n + 1], 0.5 * (g[:-1] + g[1:]), atol=0, rtol=0)


# ---------------------------------------------------------------- transforms
def _brute_to_fourier(cfg, phys):
    """Direct-sum DFT definition of fft_xz_to_fourier (independent of any FFT)."""
    nx, nz, nkx, nkz = cfg.nx, cfg.nz, cfg.nkx, cfg.nkz
    x = np.arange(nx)
    z = np.arange(nz)
    out = np.zeros((phys.shape[0], cfg.kzr, nkx + 1), np.complex128)
    kints = list(range(nkz + 1)) + [-(nkz + 1 - m) for m in range(1, nkz + 1)]
    for j in range(phys.shape[0]):
        for kk, kz in enumerate(kints):
            for i in range(nkx + 1):
                e = np.exp(-2j * np.pi * (np.outer(z * kz / nz, np.ones(nx)) + np.outer(np.ones(nz), x * i / nx)))
                out[j, kk, i] = (phys[j] * e).sum() / (nx * nz)
    return out


def _small(nx=8, ny=5, nz=8):
    return configs.small_dns(nx, ny, nz)


def test_fft_to_fourier_vs_brute_force_and_numpy():
    cfg = _small(8, 4, 8)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(1)
    phys = rng.standard_normal((cfg.ny2, cfg.nz, cfg.nx))
    f_or = o.fft_to_fourier(phys)
    f_bf = _brute_to_fourier(cfg, phys)
    f_np = layout.fft_xz_to_fourier(cfg, phys)
    assert np.abs(f_or - f_bf).max() < 1e-15
    assert np.abs(f_np - f_bf).max() < 1e-15


@pytest.mark.parametrize("nx,nz", [(16, 32), (64, 64), (32, 8)])
def test_fft_vs_numpy_and_round_trip(nx, nz):
    cfg = _small(nx, 6, nz)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(2)
    phys = rng.standard_normal((cfg.ny2, cfg.nz, cfg.nx))
    f_or = o.fft_to_fourier(phys)
    f_np = layout.fft_xz_to_fourier(cfg, phys)
    assert np.abs(f_or - f_np).max() < 1e-15 * max(1.0, np.abs(f_np).max()) * 10
    # band-limited fields survive the round trip
    b_or = o.fft_to_physical(f_or)
    b_np = layout.fft_xz_to_physical(cfg, f_np)
    assert np.abs(b_or - b_np).max() < 1e-13
    f2 = o.fft_to_fourier(b_or)
    assert np.abs(f2 - f_or).max() < 1e-15 * 10
    # the inverse is unnormalised: a unit (0,0) mode is the constant 1
    unit = np.zeros_like(f_or)
    unit[:, 0, 0] = 1.0
    assert np.abs(o.fft_to_physical(unit) - 1.0).max() < 1e-15


def test_c2r_ignores_dc_imaginary_part():
    cfg = _small(16, 4, 16)
    o = oracle.Oracle(Problem(cfg))
    f = np.zeros((cfg.ny2, cfg.kzr, cfg.kxr), np.complex128)
    f[:, 0, 0] = 1.0 + 5.0j
    assert np.abs(o.fft_to_physical(f) - 1.0).max() < 1e-15


# ---------------------------------------------------------------- Thomas solvers
def _dense(a, b, c, lo, hi):
    n = len(b)
    A = np.zeros((n, n))
    for j in range(lo, hi + 1):
        A[j, j] = b[j]
        if j > lo:
            A[j, j - 1] = a[j]
        if j < hi:
            A[j, j + 1] = c[j]
    return A


def test_thomas_real_vs_dense():
    rng = np.random.default_rng(3)
    n = 21
    a, c = rng.uniform(-1, 0, n), rng.uniform(-1, 0, n)
    b = 3.0 + rng.uniform(0, 1, n)
    g = rng.standard_normal(n)
    x = oracle.thomas_real(a, b, c, g)
    A = _dense(a, b, c, 0, n - 1)
    assert np.abs(x - np.linalg.solve(A, g)).max() < 1e-14


def test_thomas_complex_rows_1_to_ny():
    rng = np.random.default_rng(4)
    n = 19  # rows 0..ny+1 with ny = 17
    a, c = rng.uniform(0.5, 1, n), rng.uniform(0.5, 1, n)
    b = -(a + c) - rng.uniform(0.1, 1, n)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = oracle.thomas_complex(a, b, c, g)
    A = _dense(a, b, c, 1, n - 2)[1:n - 1, 1:n - 1]
    ref = np.linalg.solve(A, g[1:n - 1])
    assert np.abs(x[1:n - 1] - ref).max() < 1e-13
    assert x[0] == g[0] and x[n - 1] == g[n - 1]  # rows outside 1..Nyp untouched


def test_cn_matrix_rows_and_invariants():
    cfg = configs.small_dns(8, 17, 8)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    ny = cfg.ny
    for which in range(4):
        a, b, c, v = o.cn_matrix(which, 1)
        # interior rows: M = 1 - T*delta_y^2 is a strictly dominant M-matrix
        lo = 3 if which == 1 else 2   # a Dirichlet u2 bottom also owns row 2
        for j in range(lo, ny):
            assert a[j] < 0 and c[j] < 0 and b[j] > 1.0 and abs(b[j] - 1.0 + a[j] + c[j]) < 1e-15
    # u1: Dirichlet bottom rows 0,1; Neumann top rows Nyp, Nyp+1 (for/solver.f90:1901-1983)
    a, b, c, v = o.cn_matrix(0, 1)
    assert (a[0], b[0], c[0], v[0]) == (0, 1, 0, 0) and (a[1], b[1], c[1], v[1]) == (0, 1, 0, cfg.u_bc_ymin_c1)
    assert (a[ny], b[ny], c[ny]) == (-1, 1, 0) and v[ny] == p.met["dy"][ny] * cfg.u_bc_ymax_c1
    # u3 top Neumann carries the thermal-wind compensation value
    a, b, c, v = o.cn_matrix(2, 1)
    assert v[ny] == p.met["dy"][ny] * cfg.w_bc_ymax_c1 and cfg.w_bc_ymax_c1 == -0.5
    # u2: Dirichlet rows 1,2 bottom and Nyp,Nyp+1 top (for/solver.f90:1989-2054)
    a, b, c, v = o.cn_matrix(1, 1)
    for j in (0, 1, 2, ny, ny + 1):
        assert (a[j], b[j], c[j]) == (0, 1, 0)
    # theta: Neumann rows (0,-1,1 | dy(2) c1) and (-1,1,0 | dy(Nyp) c1)
    a, b, c, v = o.cn_matrix(3, 1)
    assert (a[1], b[1], c[1]) == (0, -1, 1) and v[1] == p.met["dy"][2] * cfg.th_bc_ymin_c1[0]
    assert (a[ny], b[ny], c[ny]) == (-1, 1, 0) and v[ny] == p.met["dy"][ny] * cfg.th_bc_ymax_c1[0]
    # a Neumann system maps a linear profile with the imposed slope onto itself
    # when T -> 0: solve and compare with the dense solution of the same rows
    rhs = np.linspace(0.0, 1.0, ny + 2)
    a, b, c, v = o.cn_matrix(3, 2)
    g = v.copy()
    g[2:ny] = rhs[2:ny]
    x = oracle.thomas_real(a, b, c, g)
    A = _dense(a, b, c, 0, ny + 1)
    assert np.abs(x - np.linalg.solve(A, g)).max() < 1e-13


# ---------------------------------------------------------------- projection
def _div(p, s):
    cfg = p.cfg
    kx, kz, dyf = p.kx[:cfg.kxr], p.kz, p.met["dyf"]
    ny = cfg.ny
    return (1j * kx[None, None, :] * s.cu1[1:ny + 1]
            + (s.cu2[2:ny + 2] - s.cu2[1:ny + 1]) / dyf[1:ny + 1, None, None]
            + 1j * kz[None, :, None] * s.cu3[1:ny + 1])


def test_rem_div_divergence_and_idempotence():
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s0 = p.initial_state(seed=5)
    o.upload_state(s0)
    o.rem_div()
    s1 = o.download_state()
    scale = np.abs(s1.cu1).max() * np.abs(p.kx).max()
    d = np.abs(_div(p, s1))
    assert d[1:-1].max() < 1e-13 * scale  # rows 2..Nyp-1 (BC rows are not forced)
    o.rem_div()
    s2 = o.download_state()
    assert np.abs(s2.cu1 - s1.cu1).max() < 1e-15 * scale * 10
    assert np.abs(s2.cu2 - s1.cu2).max() < 1e-15 * scale * 10


def test_rem_div_phi_solves_poisson_dense():
    """phi = (s0 - s1)/grad; compare with a dense solve of the BC-augmented system."""
    cfg = configs.small_dns(8, 9, 8)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s0 = p.initial_state(seed=6)
    o.upload_state(s0)
    o.rem_div()
    s1 = o.download_state()
    ny, dy, dyf = cfg.ny, p.met["dy"], p.met["dyf"]
    for (k, i) in [(0, 0), (1, 0), (0, 2), (3, 1)]:
        kx2, kz2 = p.kx[i] ** 2, p.kz[k] ** 2
        n = ny
        A = np.zeros((n, n))
        rhs = np.zeros(n, np.complex128)
        for j in range(1, ny + 1):
            r = j - 1
            if j > 1:
                A[r, r - 1] = 1.0 / (dy[j] * dyf[j])
            A[r, r] = -kx2 - kz2 - 1.0 / (dy[j + 1] * dyf[j]) - 1.0 / (dy[j] * dyf[j])
            if j < ny:
                A[r, r + 1] = 1.0 / (dy[j + 1] * dyf[j])
            rhs[r] = (1j * p.kx[i] * s0.cu1[j, k, i] + (s0.cu2[j + 1, k, i] - s0.cu2[j, k, i]) / dyf[j]
                      + 1j * p.kz[k] * s0.cu3[j, k, i])
        A[0, :] = 0
        A[0, 0] = 1
        if not (k == 0 and i == 0):
            A[0, 1] = -1
        rhs[0] = 0
        A[n - 1, :] = 0
        A[n - 1, n - 2], A[n - 1, n - 1] = 1, -1
        rhs[n - 1] = 0
        phi = np.linalg.solve(A, rhs)
        # velocity correction cu2(j) -= (phi(j)-phi(j-1))/dy(j) for j = 2..Nyp
        du2 = s0.cu2[2:ny + 1, k, i] - s1.cu2[2:ny + 1, k, i]
        ref = (phi[1:] - phi[:-1]) / dy[2:ny + 1]
        assert np.abs(du2 - ref).max() < 1e-12 * max(1e-30, np.abs(ref).max()) + 1e-20
        if k == 0 and i == 0:
            assert abs(phi[0]) <= 1e-14 * max(1e-30, np.abs(phi).max())  # phi(1) = 0 row


# ---------------------------------------------------------------- RKW3 closed forms
def _linear_cfg(nx=8, ny=9, nz=8, **kw):
    base = dict(nx=nx, ny=ny, nz=nz, n_th=1, flavor="None", re=1e300, nu_v_scale=0.0,
                ro=1e300, delta=1.0, grav_y=0.0, ic_type=4, u_bc_ymin=NEUMANN, u_bc_ymax=NEUMANN,
                w_bc_ymin=NEUMANN, w_bc_ymax=NEUMANN, v_bc_ymin=DIRICHLET, v_bc_ymax=DIRICHLET,
                th_bc_ymin=[NEUMANN], th_bc_ymax=[NEUMANN], th_bc_ymin_c1=[0.0], th_bc_ymax_c1=[0.0],
                dthdx=[0.0], dthdz=[0.0], bc_compensation=False, delta_t=0.01)
    base.update(kw)
    return Config(**base).derive()


def _P(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


def _zero_state(cfg):
    z = np.zeros((cfg.ny2, cfg.kzr, cfg.kxr), np.complex128)
    return FourierState(z.copy(), z.copy(), z.copy(), z.copy(), [z.copy() for _ in range(cfg.n_th)])


def test_rkw3_horizontal_diffusion_polynomial():
    cfg = _linear_cfg(re=50.0, delta_t=0.05)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s = _zero_state(cfg)
    # kz = +-1, kx = 0 (a real field needs the conjugate pair): divergence-free and
    # free of nonlinear terms
    s.cu1[:, 1, 0] = 0.3 + 0.1j
    s.cu1[:, cfg.twonkz, 0] = 0.3 - 0.1j
    o.upload_state(s)
    o.step(time_step=1)
    r = o.download_state()
    z = -cfg.nu * p.kz[1] ** 2 * cfg.delta_t
    ny = cfg.ny
    expect = (0.3 + 0.1j) * _P(z)
    assert np.abs(r.cu1[1:ny + 1, 1, 0] - expect).max() < 1e-15


def test_rkw3_coriolis_rotation_polynomial():
    cfg = _linear_cfg(ro=1.0 / 1.3, delta_t=0.1)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s = _zero_state(cfg)
    s.cu1[:, 0, 0] = 0.7
    s.cu3[:, 0, 0] = -0.2
    o.upload_state(s)
    o.step(time_step=1)
    r = o.download_state()
    f = cfg.ro_inv / cfg.delta
    w = (0.7 - 0.2j) * _P(-1j * f * cfg.delta_t)
    ny = cfg.ny
    assert np.abs(r.cu1[1:ny + 1, 0, 0] - w.real).max() < 1e-15
    assert np.abs(r.cu3[1:ny + 1, 0, 0] - w.imag).max() < 1e-15


def test_thermal_wind_advection_phase():
    ro_inv = 1e-10
    cfg = _linear_cfg(flavor="Front", ro=1.0 / ro_inv, delta=1.0, dthdx=[ro_inv], delta_t=0.02)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s = _zero_state(cfg)
    s.cu1[:, 1, 0] = 1.0
    s.cu1[:, cfg.twonkz, 0] = 1.0
    o.upload_state(s)
    o.step(time_step=1)
    r = o.download_state()
    ny, gyf = cfg.ny, p.met["gyf"]
    vg = cfg.dthdx[0] * (gyf - 0.5 * cfg.ly) * cfg.delta / cfg.ro_inv
    expect = _P(-1j * p.kz[1] * vg * cfg.delta_t)
    assert np.abs(r.cu1[2:ny, 1, 0] - expect[2:ny]).max() < 1e-12


# ---------------------------------------------------------------- equivariance
def _shift(cfg, s, sx, sz):
    i = np.arange(cfg.kxr)
    nkz = cfg.nkz
    kint = np.array(list(range(nkz + 1)) + [-(nkz + 1 - m) for m in range(1, nkz + 1)])
    ph = np.exp(-2j * np.pi * (i[None, :] * sx / cfg.nx + kint[:, None] * sz / cfg.nz))[None]
    return FourierState(s.cu1 * ph, s.cu2 * ph, s.cu3 * ph, s.cp * ph, [t * ph for t in s.cth])


def test_substep_is_translation_equivariant():
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    s0 = p.initial_state(seed=7, theta_noise=1e-3)
    # larger perturbations so the nonlinear terms matter
    for a in (s0.cu1, s0.cu2, s0.cu3):
        a *= 300.0
    o.upload_state(s0)
    o.rem_div()
    s0 = o.download_state()
    o.step(time_step=1)
    a = _shift(cfg, o.download_state(), 3, 5)
    o.upload_state(_shift(cfg, s0, 3, 5))
    o.step(time_step=1)
    b = o.download_state()
    for x, y in ((a.cu1, b.cu1), (a.cu2, b.cu2), (a.cu3, b.cu3), (a.cp, b.cp), (a.cth[0], b.cth[0])):
        assert np.abs(x - y).max() < 1e-13 * max(1.0, np.abs(x).max())


def test_ten_steps_stay_divergence_free_and_finite():
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    o.upload_state(p.initial_state(seed=8, theta_noise=1e-3))
    o.rem_div()
    for t in range(10):
        o.step(time_step=t + 1)
    s = o.download_state()
    assert all(np.isfinite(a).all() for a in (s.cu1, s.cu2, s.cu3, s.cp, s.cth[0]))
    d = np.abs(_div(p, s))
    assert d[1:-1].max() < 1e-13 * max(1e-30, np.abs(s.cu1).max())


# ---------------------------------------------------------------- courant
def test_courant_limits_closed_form():
    """courant (for/flow.f90:746-806) on a state whose extrema are known: u1 = A sin(2 pi x / Lx)
    plus the thermal wind the 'Front' flavour adds to u3."""
    cfg = configs.small_dns(32, 17, 16)
    p = Problem(cfg)
    o = oracle.Oracle(p, threads=2)
    shp = (cfg.ny2, cfg.nz, cfg.nx)
    x = np.arange(cfg.nx) * cfg.lx / cfg.nx
    zero = np.zeros(shp)
    pi32 = 3.1415927410125732
    dx, dz = cfg.lx / cfg.nx, cfg.lz / cfg.nz
    f = cfg.ro_inv / cfg.delta
    gyf = p.met["gyf"]
    vg = np.abs((1.0 / f) * cfg.dthdx[0] * (gyf[cfg.jstart:cfg.jend + 1] - 0.5 * cfg.ly)).max()
    limits = [1.0, 0.5 * min(dx, dz) ** (2 * cfg.beta) / cfg.nu, 2 * pi32 / abs(f) / 20,
              float(np.float32(0.1)) * 2 * pi32 / np.sqrt(abs(cfg.th_bc_ymin_c1[0]))]
    for amp in (1e-3, 5.0, 40.0):
        u1 = np.broadcast_to(amp * np.sin(2 * np.pi * x / cfg.lx), shp).copy()
        s = FourierState(layout.fft_xz_to_fourier(cfg, u1), layout.compact_zeros(cfg),
                         layout.compact_zeros(cfg), layout.compact_zeros(cfg), [layout.compact_zeros(cfg)])
        o.upload_state(s)
        cfl = 0.5
        got = o.courant_dt_state(cfl)
        ref = min(limits + [cfl * dx / amp, cfl * dz / vg])
        assert abs(got - ref) <= 1e-12 * ref, (amp, got, ref)


def test_courant_schedule_cases():
    """The event scheduling of courant (for/flow.f90:810-851)."""
    sch = oracle.lib().oracle_courant_schedule
    r_max = float(np.float32(1.0) - np.float32(0.025))
    # far from any event: dt itself
    assert sch(1e-3, 0.0, 10.0, 5.0, 5.0) == 1e-3
    # the next event is closer than dt: go straight there
    assert sch(1e-3, 4.9995, 10.0, 5.0, 5.0) == (5.0 - 4.9995) + 1e-14
    # between one and two steps away: two equal steps
    nxt = 1.5e-3
    assert sch(1e-3, 5.0 - nxt, 10.0, 5.0, 5.0) == 1.0 / (1.0 + r_max) * (5.0 - (5.0 - nxt))
    # the caps are overridden by the scheduling branch, as in the reference
    assert sch(2000.0, 0.0, 1e9, 1e9, 1e9) == 2000.0
