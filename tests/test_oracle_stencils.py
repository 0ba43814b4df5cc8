"""Closed-form pins for the oracle's wall-normal (y) stencils.

The oracle restates rk_chan_1's staggered-grid stencils: fields u1, u3, theta live
at GYF points, u2 at GY points, with dy(j) = GYF(j) - GYF(j-1) and
dyf(j) = GY(j+1) - GY(j).  The GPU path is held to the oracle, so each stencil is
checked here against what the mathematics of that grid fixes, with no copy of the
stencil in the test:

* the buoyancy's GYF -> GY interpolation is linear interpolation, exact for
  profiles linear in y (for/solver.f90:518-538);
* the u1*u2 and u3*u2 products take u1, u3 to GY as the width-weighted mean of
  the two cells that meet at GY(j) (for/solver.f90:700-730): for a profile linear
  in y that is its value at the centre of the double cell,
  (GY(j-1) + GY(j+1))/2, not at GY(j);
* conservative flux differences telescope: sum_j dyf_j * d(F)/dy_j equals the
  difference of the end fluxes (for/solver.f90:740-790, 950-964), and with one
  factor constant the flux difference is that factor times the other's difference;
* the GY-point d(u2 u2)/dy of a profile linear in y is exact: the centred
  difference of a quadratic over [GYF(j-1), GYF(j)] is its derivative at the midpoint
  (for/solver.f90:795-815);
* the vertical viscous operator (explicit half for/solver.f90:850-903, 967-1002;
  implicit rows 1013-1045, 1072-1105, 1127-1161, 1181-1217) is a three-point
  finite-volume operator with two-point face gradients, so it is fixed by its
  action on 1, y and y^2: zero, zero, and the face-gradient difference
  (2 ymid(j+1) - 2 ymid(j)) / cell width;  with an eddy viscosity the face flux
  carries nu_t at the face (GY for GYF fields, the GYF average for u2);
* the BC posts reproduce the imposed wall values and wall gradients of the mean
  mode across the first interior face (for/flow.f90:1161-1360, 1596-1689).

A dy <-> dyf exchange in any one of these stencils fails at least one test here.
"""
import numpy as np

import oracle
from diablo3_amd import layout
from diablo3_amd.config import Config, DIRICHLET, NEUMANN
from diablo3_amd.ic import FourierState
from diablo3_amd.problem import Problem

INF = float("inf")


def _cfg(nx=8, ny=17, nz=8, **kw):
    """A channel with every term off unless a test switches it on: no viscosity,
    no rotation, no gravity, no thermal wind, a stretched y grid (CS 1.75)."""
    base = dict(nx=nx, ny=ny, nz=nz, n_th=1, flavor="None", re=1e300, nu_v_scale=1.0,
                ro=INF, delta=1.0, grav_x=0.0, grav_y=0.0, grav_z=0.0, ic_type=4,
                u_bc_ymin=DIRICHLET, u_bc_ymax=DIRICHLET, w_bc_ymin=DIRICHLET, w_bc_ymax=DIRICHLET,
                v_bc_ymin=DIRICHLET, v_bc_ymax=DIRICHLET,
                th_bc_ymin=[DIRICHLET], th_bc_ymax=[DIRICHLET], th_bc_ymin_c1=[0.0],
                th_bc_ymax_c1=[0.0], pr=[1.0], ri=[0.0], dthdx=[0.0], dthdz=[0.0],
                bc_compensation=False, delta_t=0.05, grid_cs=1.75)
    base.update(kw)
    return Config(**base).derive()


def _zero(cfg):
    z = layout.compact_zeros(cfg)
    return FourierState(z.copy(), z.copy(), z.copy(), z.copy(), [z.copy() for _ in range(cfg.n_th)])


def _set_kz(a, cfg, k, prof):
    """A real field with the z mode +k (profile prof) and its conjugate -k."""
    a[:, k, 0] = prof
    a[:, cfg.twonkz + 1 - k, 0] = np.conj(prof)


def _substep(cfg, s, rk_step=1, **kw):
    p = Problem(cfg)
    o = oracle.Oracle(p)
    o.upload_state(s)
    o.rk_substep(rk_step, **kw)
    return p, o


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ----------------------------------------------------------------- A7 buoyancy
def test_buoyancy_interpolation_exact_for_linear_theta():
    """cf2 += g_y theta(GY_j): a theta linear in y interpolates exactly; the (0,0)
    mode gets no buoyancy (for/solver.f90:518-538)."""
    g = 0.7
    cfg = _cfg(grav_y=g)
    s = _zero(cfg)
    p = Problem(cfg)
    gy, gyf, ny = p.met["gy"], p.met["gyf"], cfg.ny
    a, b = 0.3 - 0.2j, 1.1 + 0.4j
    s.cth[0][:, 0, 1] = a + b * gyf                    # kx mode 1
    _set_kz(s.cth[0], cfg, 1, -0.5 + 2.0j * gyf)       # kz mode +-1
    s.cth[0][:, 0, 0] = np.random.default_rng(0).standard_normal(cfg.ny2)  # the mean
    p, o = _substep(cfg, s)
    cf, _ = o.download_rk()
    j = np.arange(2, ny + 1)
    assert _rel(cf[1][j, 0, 1], g * (a + b * gy[j])) < 1e-14
    assert _rel(cf[1][j, 1, 0], g * (-0.5 + 2.0j * gy[j])) < 1e-14
    assert np.all(cf[1][j, 0, 0] == 0.0)


# ----------------------------------------------------------------- A9 products
def test_gy_products_two_cell_mean_for_linear_u1_u3():
    """u1*u2 and u3*u2 at GY: u1, u3 go to GY_j as the mean over the two cells
    [GY_{j-1}, GY_{j+1}] of their cell values (width-weighted by dyf), which for a
    profile linear in y is its value at (GY_{j-1} + GY_{j+1})/2; so
    cf2 = -i kx u2 u1(ym) - i kz u2 u3(ym) (for/solver.f90:700-730)."""
    cfg = _cfg()
    s = _zero(cfg)
    p = Problem(cfg)
    gy, gyf, ny = p.met["gy"], p.met["gyf"], cfg.ny
    c = 0.5 + np.random.default_rng(1).standard_normal(cfg.ny2)  # u2 mean at GY
    s.cu2[:, 0, 0] = c
    a, b = 0.2 + 0.1j, -0.8 + 0.3j
    s.cu1[:, 0, 1] = a + b * gyf
    _set_kz(s.cu3, cfg, 1, 0.4 - 1.3j * gyf)
    p, o = _substep(cfg, s)
    cf, _ = o.download_rk()
    j = np.arange(2, ny + 1)
    ym = 0.5 * (gy[j - 1] + gy[j + 1])
    assert _rel(cf[1][j, 0, 1], -1j * p.kx[1] * c[j] * (a + b * ym)) < 1e-14
    assert _rel(cf[1][j, 1, 0], -1j * p.kz[1] * c[j] * (0.4 - 1.3j * ym)) < 1e-14
    # constants pass unchanged
    s.cu1[:, 0, 1] = a
    p, o = _substep(cfg, s)
    cf, _ = o.download_rk()
    assert _rel(cf[1][j, 0, 1], -1j * p.kx[1] * c[j] * a) < 1e-14


def _flux_state(cfg, A, B, T, c):
    s = _zero(cfg)
    s.cu2[:, 0, 0] = c        # u2 mean at GY
    s.cu1[:, 0, 1] = A        # u1 kx mode 1 at GYF
    _set_kz(s.cu3, cfg, 1, B)  # u3 kz mode 1 at GYF
    s.cth[0][:, 0, 2] = T     # theta kx mode 2 at GYF
    return s


def test_vertical_flux_differences_telescope():
    """d(u1 u2)/dy, d(u3 u2)/dy, d(theta u2)/dy at GYF are conservative: with the
    face flux F_j = u2(GY_j) times the mean of the two neighbouring GYF values,
    sum_{j} dyf_j cf_j = -(F_{jend+1} - F_{jstart}) (for/solver.f90:740-790, 950-964)."""
    cfg = _cfg()
    rng = np.random.default_rng(2)
    n2 = cfg.ny2
    A, B, T = (rng.standard_normal(n2) + 1j * rng.standard_normal(n2) for _ in range(3))
    c = rng.standard_normal(n2)
    p, o = _substep(cfg, _flux_state(cfg, A, B, T, c))
    cf, cfth = o.download_rk()
    dyf = p.met["dyf"]
    for got, prof, js, je in ((cf[0][:, 0, 1], A, cfg.jstart, cfg.jend),
                              (cf[2][:, 1, 0], B, cfg.jstart, cfg.jend),
                              (cfth[0][:, 0, 2], T, cfg.jstart_th[0], cfg.jend_th[0])):
        F = np.zeros(n2, np.complex128)
        F[1:] = 0.5 * (prof[1:] + prof[:-1]) * c[1:]
        j = np.arange(js, je + 1)
        lhs = (dyf[j] * got[j]).sum()
        assert abs(lhs + (F[je + 1] - F[js])) < 1e-13 * np.abs(F).max()


def test_vertical_flux_with_uniform_factor():
    """With u1 (u3, theta) uniform in y, the flux difference is that value times
    the GYF-cell difference of u2: cf = -A (u2_{j+1} - u2_j) / dyf_j."""
    cfg = _cfg()
    n2 = cfg.ny2
    c = np.random.default_rng(3).standard_normal(n2)
    A, B, T = (np.full(n2, v) for v in (0.7 - 0.1j, -0.3 + 0.9j, 1.2 + 0.2j))
    p, o = _substep(cfg, _flux_state(cfg, A, B, T, c))
    cf, cfth = o.download_rk()
    dyf = p.met["dyf"]
    for got, v, js, je in ((cf[0][:, 0, 1], A[0], cfg.jstart, cfg.jend),
                           (cf[2][:, 1, 0], B[0], cfg.jstart, cfg.jend),
                           (cfth[0][:, 0, 2], T[0], cfg.jstart_th[0], cfg.jend_th[0])):
        j = np.arange(js, je + 1)
        assert _rel(got[j], -v * (c[j + 1] - c[j]) / dyf[j]) < 1e-13


def test_u2u2_derivative_exact_for_linear_u2():
    """d(u2 u2)/dy at GY from the GYF means of u2: for u2 = a + b y the GYF mean is
    u2(GYF) exactly and the difference of squares over dy_j is the derivative at the
    midpoint of GYF(j-1), GYF(j): cf2 = -b (2a + b (GYF_j + GYF_{j-1}))
    (for/solver.f90:795-815); for any profile the dy-weighted sum telescopes."""
    cfg = _cfg()
    p = Problem(cfg)
    gy, gyf, dy, ny = p.met["gy"], p.met["gyf"], p.met["dy"], cfg.ny
    a, b = 0.4, -1.7
    s = _zero(cfg)
    s.cu2[:, 0, 0] = a + b * gy
    p, o = _substep(cfg, s)
    cf, _ = o.download_rk()
    j = np.arange(2, ny + 1)
    assert _rel(cf[1][j, 0, 0], -b * (2 * a + b * (gyf[j] + gyf[j - 1]))) < 1e-14
    # any profile: sum_{j=2}^{ny} dy_j cf2_j = -(U_ny^2 - U_1^2), U = GYF mean of u2
    c = np.random.default_rng(4).standard_normal(cfg.ny2)
    s = _zero(cfg)
    s.cu2[:, 0, 0] = c
    p, o = _substep(cfg, s)
    cf, _ = o.download_rk()
    U = 0.5 * (c[:-1] + c[1:])          # U[j] at GYF_j
    lhs = (dy[j] * cf[1][j, 0, 0]).sum()
    assert abs(lhs + (U[ny] ** 2 - U[1] ** 2)) < 1e-13 * max(1.0, np.abs(U).max() ** 2)


# ----------------------------------------------------------------- A10/A11 viscous operator
def _lap_expected(centres, widths, coef, j):
    """(1, y, y^2) -> the FV operator's values at rows j: the face gradient of y^2
    between centres m and m+1 is exactly centres[m] + centres[m+1], so the operator
    maps coef = (c0, c1, c2) to c2 * (y_{j+1} - y_{j-1}) / width_j."""
    return coef[2] * (centres[j + 1] - centres[j - 1]) / widths[j]


def test_explicit_viscous_half_on_quadratics():
    """The explicit CN half r += (nu h/2) nu_v L(u) with L the FV second derivative:
    L(1) = L(y) = 0 and L(y^2) = (y_{j+1} - y_{j-1}) / cell width, at GYF for u1,
    u3, theta (cells of width dyf, /Pr for theta) and at GY for u2 (cells of width
    dy) (for/solver.f90:850-870, 967-980)."""
    dt, re, nvs, pr = 0.1, 10.0, 0.7, 2.0
    cfg = _cfg(re=re, nu_v_scale=nvs, pr=[pr], delta_t=dt)
    p = Problem(cfg)
    gy, gyf, dy, dyf, ny = p.met["gy"], p.met["gyf"], p.met["dy"], p.met["dyf"], cfg.ny
    temp1 = 0.5 * cfg.nu * dt * 8.0 / 15.0
    cq = {0: (0.3, -0.4, 1.3), 2: (-0.2, 0.9, -0.6), 3: (0.1, 0.2, 0.8)}
    c2 = (0.05, 0.3, 0.45)
    s = _zero(cfg)
    for which, arr in ((0, s.cu1), (2, s.cu3), (3, s.cth[0])):
        co = cq[which]
        arr[:, 0, 0] = co[0] + co[1] * gyf + co[2] * gyf ** 2
    s.cu2[:, 0, 0] = c2[0] + c2[1] * gy + c2[2] * gy ** 2
    p, o = _substep(cfg, s)
    cf, cfth = o.download_rk()
    h = dt * 8.0 / 15.0
    # r = u + h cf + T L(u) on the rows each field is advanced on; everything is a
    # mean mode, so cf (the advection of the mean profiles) is its own physical value
    for which, cfw in ((0, cf[0]), (2, cf[2]), (3, cfth[0])):
        co = cq[which]
        js, je = (cfg.jstart_th[0], cfg.jend_th[0]) if which == 3 else (cfg.jstart, cfg.jend)
        j = np.arange(js, je + 1)
        r = o.debug_rhs(which)[j, 0, 0]
        q = co[0] + co[1] * gyf[j] + co[2] * gyf[j] ** 2
        t = temp1 * nvs / (pr if which == 3 else 1.0)
        expect = t * _lap_expected(gyf, dyf, co, j)
        got = r - q - h * cfw[j, 0, 0].real
        assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max(), which
    j = np.arange(2, ny + 1)
    r2 = o.debug_rhs(1)[j, 0, 0]
    q2 = c2[0] + c2[1] * gy[j] + c2[2] * gy[j] ** 2
    expect = temp1 * nvs * _lap_expected(gy, dy, c2, j)
    got = r2 - q2 - h * cf[1][j, 0, 0].real
    assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


def _apply(a, b, c, q, j):
    return a[j] * q[j - 1] + b[j] * q[j] + c[j] * q[j + 1]


def test_cn_rows_on_quadratics():
    """Interior rows of every CN system are 1 - T L with the same FV operator:
    M(1) = 1, M(y) = y, M(y^2) = y^2 - T (y_{j+1} - y_{j-1}) / width_j
    (for/solver.f90:1020-1045, 1081-1105, 1136-1161, 1192-1217).  Three conditions
    per row fix all three coefficients."""
    dt, re, nvs, pr = 0.1, 10.0, 0.7, 2.0
    cfg = _cfg(re=re, nu_v_scale=nvs, pr=[pr], delta_t=dt)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    gy, gyf, dy, dyf, ny = p.met["gy"], p.met["gyf"], p.met["dy"], p.met["dyf"], cfg.ny
    for rk, hfac in ((1, 8.0 / 15.0), (3, 5.0 / 15.0)):
        temp1 = 0.5 * cfg.nu * dt * hfac
        for which in range(4):
            a, b, c, _ = o.cn_matrix(which, rk)
            yc, w = (gy, dy) if which == 1 else (gyf, dyf)
            j = np.arange(3, ny - 1) if which == 1 else np.arange(2, ny)
            T = temp1 * nvs / (pr if which == 3 else 1.0)
            one = np.ones(cfg.ny2)
            assert np.abs(_apply(a, b, c, one, j) - 1.0).max() < 1e-15
            assert np.abs(_apply(a, b, c, yc, j) - yc[j]).max() < 1e-14
            expect = yc[j] ** 2 - T * _lap_expected(yc, w, (0, 0, 1.0), j)
            assert np.abs(_apply(a, b, c, yc ** 2, j) - expect).max() < 1e-14, (which, rk)


def _les_cfg(**kw):
    base = dict(use_les=True, les_model_type=5, re=10.0, nu_v_scale=0.7, pr=[2.0], delta_t=0.1)
    base.update(kw)
    return _cfg(**base)


def test_cn_rows_with_eddy_viscosity():
    """With nu_t the face flux carries nu_t at the face: at GY for GYF fields, the
    mean of the two GY values at GYF for u2 (for/solver.f90:1030-1043, 1090-1103).
    The operator on 1, y, y^2 is then 0, the face difference of nu_t over the cell,
    and the face difference of nu_t (y_m + y_{m+1})."""
    cfg = _les_cfg()
    p = Problem(cfg)
    o = oracle.Oracle(p)
    gy, gyf, dy, dyf, ny = p.met["gy"], p.met["gyf"], p.met["dy"], p.met["dyf"], cfg.ny
    nut = 0.01 + 0.02 * np.random.default_rng(5).random(cfg.ny2)
    dt, nvs, pr = cfg.delta_t, cfg.nu_v_scale, cfg.pr[0]
    rk = 2
    h = dt * 2.0 / 15.0
    temp1, temp2 = 0.5 * cfg.nu * h, 0.5 * h
    for which in range(4):
        a, b, c, _ = o.cn_matrix_nut(which, rk, nut)
        if which == 1:
            yc, w, j = gy, dy, np.arange(3, ny - 1)
            nf = np.zeros(cfg.ny2)
            nf[:-1] = 0.5 * (nut[:-1] + nut[1:])      # nu_t at GYF_m, face between GY_m, GY_m+1
            face = lambda m: nf[m]                    # face above centre m
        else:
            yc, w, j = gyf, dyf, np.arange(2, ny)
            face = lambda m: nut[m + 1]              # GY_{m+1} is the face above GYF_m
        T = temp1 * nvs / (pr if which == 3 else 1.0)
        one = np.ones(cfg.ny2)
        assert np.abs(_apply(a, b, c, one, j) - 1.0).max() < 1e-15, which
        exp_y = yc[j] - temp2 * (face(j) - face(j - 1)) / w[j]
        assert np.abs(_apply(a, b, c, yc, j) - exp_y).max() < 1e-14, which
        exp_y2 = (yc[j] ** 2 - T * (yc[j + 1] - yc[j - 1]) / w[j]
                  - temp2 * (face(j) * (yc[j] + yc[j + 1]) - face(j - 1) * (yc[j - 1] + yc[j])) / w[j])
        assert np.abs(_apply(a, b, c, yc ** 2, j) - exp_y2).max() < 1e-14, which


def test_explicit_eddy_viscosity_half_on_linear_profiles():
    """The explicit nu_t half r += (h/2) delta_y(nu_t delta_y u) (for/solver.f90:875-903,
    983-1002) on a u1 (theta) linear in y with slope b: the face gradient is b, so the
    term is (h/2) b (nu_t(GY_{j+1}) - nu_t(GY_j)) / dyf_j at every (x, z); nu_t and
    kappa_t come from the AMD closure of a 3D perturbation carried by u2 and u3 alone."""
    cfg = _les_cfg(nx=16, nz=16, ny=17, u_bc_ymin_c1=0.0, th_bc_ymin=[DIRICHLET],
                   th_bc_ymax=[DIRICHLET])
    p = Problem(cfg)
    gy, gyf, dyf, ny = p.met["gy"], p.met["gyf"], p.met["dyf"], cfg.ny
    b1, a1, bt, at = 0.8, -0.2, -0.6, 0.3
    cfg = cfg.copy(u_bc_ymin_c1=a1 + b1 * gyf[1], u_bc_ymax_c1=a1 + b1 * gyf[ny],
                   th_bc_ymin_c1=[at + bt * gyf[1]], th_bc_ymax_c1=[at + bt * gyf[ny]])
    p = Problem(cfg)
    s = _zero(cfg)
    s.cu1[:, 0, 0] = a1 + b1 * gyf
    s.cth[0][:, 0, 0] = at + bt * gyf
    rng = np.random.default_rng(6)
    prof = (rng.standard_normal(cfg.ny2) + 1j * rng.standard_normal(cfg.ny2)) * 0.3
    prof2 = (rng.standard_normal(cfg.ny2) + 1j * rng.standard_normal(cfg.ny2)) * 0.3
    for (k, i) in ((1, 1), (2, 1), (1, 2)):
        s.cu3[:, k, i] = prof * (1 + k + i)
        s.cu2[:, k, i] = prof2 * (2 - k + i)
    o = oracle.Oracle(p)
    o.upload_state(s)
    o.rk_substep(1, create_new_flow=0)
    nut, kappa = o.download_les()
    assert nut[2:ny].max() > 1e-6 and kappa[0][2:ny].max() > 1e-6  # the closure is active
    cf, cfth = o.download_rk()
    h = cfg.delta_t * 8.0 / 15.0
    j = np.arange(2, ny)
    for which, a0, b0, kk, cfj in ((0, a1, b1, nut, cf[0]), (3, at, bt, kappa[0], cfth[0])):
        r = o.debug_rhs(which)
        base = a0 + b0 * gyf[j]
        # the explicit part of cr: the state (mean only) plus h cf, in physical space
        crp = layout.fft_xz_to_physical(cfg, cfj)[j] * h + base[:, None, None]
        got = r[j] - crp
        expect = 0.5 * h * b0 * (kk[j + 1] - kk[j]) / dyf[j][:, None, None]
        assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max() + 1e-17, which


def test_couette_profile_is_steady():
    """A Couette profile u1 = a + b y (and theta likewise) between Dirichlet walls
    holding its wall values stays put through full RKW3 steps: the vertical operator
    annihilates linear profiles and nothing else acts on the mean mode."""
    cfg = _cfg(re=5.0, delta_t=0.05)
    p = Problem(cfg)
    gyf, ny = p.met["gyf"], cfg.ny
    a, b = 0.25, 0.9
    cfg = cfg.copy(u_bc_ymin_c1=a + b * gyf[1], u_bc_ymax_c1=a + b * gyf[ny],
                   th_bc_ymin_c1=[-b * gyf[1]], th_bc_ymax_c1=[-b * gyf[ny]])
    p = Problem(cfg)
    s = _zero(cfg)
    s.cu1[:, 0, 0] = a + b * gyf
    s.cth[0][:, 0, 0] = -b * gyf
    o = oracle.Oracle(p)
    o.upload_state(s)
    for t in range(3):
        o.step(time_step=t + 1)
    r = o.download_state()
    j = np.arange(1, ny + 1)
    assert np.abs(r.cu1[j, 0, 0] - (a + b * gyf[j])).max() < 1e-14
    assert np.abs(r.cth[0][j, 0, 0] - (-b * gyf[j])).max() < 1e-14


# ----------------------------------------------------------------- A17 BC posts
def _posted(cfg, seed=7, time=0.0):
    p = Problem(cfg)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(seed)
    s = _zero(cfg)
    for arr in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        arr[:] = rng.standard_normal(arr.shape) + 1j * rng.standard_normal(arr.shape)
        arr[:, 0, 0] = arr[:, 0, 0].real
    o.upload_state(s)
    o.bc_post(time)
    return p, o.download_state()


def test_bc_post_neumann_wall_gradients():
    """Neumann posts: the mean mode's gradient across the first interior face is the
    imposed c1 -- (a2 - a1)/dy_2 at the bottom and (a_ny - a_{ny-1})/dy_ny at the top
    for GYF fields, (u2_2 - u2_1)/dyf_1 at the bottom for u2 -- and every other mode
    has zero gradient there (for/flow.f90:1184-1195, 1212-1220, 1233-1242, 1276-1288,
    1318-1333, 1344-1356, 1605-1612, 1640-1648)."""
    cfg = _cfg(u_bc_ymin=NEUMANN, u_bc_ymax=NEUMANN, w_bc_ymin=NEUMANN, w_bc_ymax=NEUMANN,
               v_bc_ymin=NEUMANN, v_bc_ymax=NEUMANN, th_bc_ymin=[NEUMANN], th_bc_ymax=[NEUMANN],
               u_bc_ymin_c1=0.3, u_bc_ymax_c1=-0.7, w_bc_ymin_c1=1.1, w_bc_ymax_c1=0.4,
               v_bc_ymin_c1=-0.9, v_bc_ymax_c1=0.6, th_bc_ymin_c1=[0.25], th_bc_ymax_c1=[-1.5])
    p, s = _posted(cfg)
    dy, dyf, ny = p.met["dy"], p.met["dyf"], cfg.ny
    for arr, lo, hi in ((s.cu1, 0.3, -0.7), (s.cu3, 1.1, 0.4), (s.cth[0], 0.25, -1.5)):
        m = arr[:, 0, 0]
        assert abs((m[2] - m[1]) / dy[2] - lo) < 1e-13
        assert abs((m[1] - m[0]) / dy[1] - lo) < 1e-13
        assert abs((m[ny] - m[ny - 1]) / dy[ny] - hi) < 1e-13
        # the second ghost repeats the same step (dy(Nyp), for/flow.f90:1285-1286)
        assert abs((m[ny + 1] - m[ny]) / dy[ny] - hi) < 1e-13
        rest = arr.copy()
        rest[:, 0, 0] = 0
        assert np.array_equal(rest[1], rest[2]) and np.array_equal(rest[0], rest[1])
        assert np.array_equal(rest[ny], rest[ny - 1]) and np.array_equal(rest[ny + 1], rest[ny])
    m = s.cu2[:, 0, 0]
    assert abs((m[2] - m[1]) / dyf[1] - (-0.9)) < 1e-13
    # the u2 top ghost steps by dy(Nyp) (for/flow.f90:1352)
    assert abs((m[ny + 1] - m[ny]) / dy[ny] - 0.6) < 1e-13


def test_bc_post_dirichlet_wall_values():
    """Dirichlet posts: GYF fields hold c1 on both wall planes of the mean mode and 0
    elsewhere (the top u3 ghost is zeroed, for/flow.f90:1311); u2's wall value is the
    mean of the two planes around the wall, GYF(1) = (GY(1) + GY(2))/2, so
    (u2_1 + u2_2)/2 = c1 -- the bottom post applies it to every mode, the top one to the
    mean only (for/flow.f90:1168-1181, 1222-1232, 1253-1264, 1336-1342)."""
    cfg = _cfg(u_bc_ymin_c1=0.3, u_bc_ymax_c1=-0.7, w_bc_ymin_c1=1.1, w_bc_ymax_c1=0.4,
               v_bc_ymin_c1=-0.9, v_bc_ymax_c1=0.6, th_bc_ymin_c1=[0.25], th_bc_ymax_c1=[-1.5])
    p, s = _posted(cfg)
    ny = cfg.ny
    for arr, lo, hi, ghost_hi in ((s.cu1, 0.3, -0.7, -0.7), (s.cu3, 1.1, 0.4, 0.0),
                                  (s.cth[0], 0.25, -1.5, -1.5)):
        assert arr[0, 0, 0] == lo and arr[1, 0, 0] == lo
        assert arr[ny, 0, 0] == hi and arr[ny + 1, 0, 0] == ghost_hi
        for jj in (0, 1, ny, ny + 1):
            plane = arr[jj].copy()
            plane[0, 0] = 0
            assert np.all(plane == 0)
    assert np.abs(0.5 * (s.cu2[1] + s.cu2[2]) - (-0.9)).max() < 1e-15
    assert abs(0.5 * (s.cu2[ny, 0, 0] + s.cu2[ny + 1, 0, 0]) - 0.6) < 1e-15
