"""Pins of the filtered-DNS training data (diablo3_amd/sgs_data.py) by closed
forms: the discrete top-hat's transfer function on single modes, cell averages
of piecewise-constant profiles, and the subgrid stress of a single x mode,
tau_11 = 1/2 + H(2k)/2 cos 2kx - H(k)^2 cos^2 kx, which vanishes for a constant."""
import numpy as np

from diablo3_amd import configs, grid, sgs_data
from diablo3_amd.problem import Problem


def H(k, dx, fx):
    m = np.arange(-fx // 2, fx // 2 + 1)
    w = np.ones(fx + 1)
    w[0] = w[-1] = 0.5
    return np.sum(w / fx * np.cos(k * m * dx))


def test_tophat_transfer_function():
    nx, nz, fx = 64, 32, 4
    lx = 2 * np.pi * 3
    x = np.arange(nx) * lx / nx
    for m in (0, 1, 5, 9):
        k = 2 * np.pi * m / lx
        f = np.broadcast_to(np.cos(k * x + 0.3), (3, nz, nx)).copy()
        g = sgs_data.tophat_xz(f, fx, 2)
        want = H(k, lx / nx, fx) * np.cos(k * x[::fx] + 0.3)
        assert np.allclose(g[1, 0], want, atol=1e-13), m


def test_overlap_rows_average_piecewise_constant():
    dns = grid.metrics(grid.create_grid(64, 2.0, 1.75))
    les = grid.metrics(grid.create_grid(32, 2.0, 1.75))
    yf = sgs_data.y_filters(dns, les)
    rs = yf.gyf[1:33].sum(axis=1)
    assert np.allclose(rs, 1.0)
    assert np.allclose(yf.gy[2:32].sum(axis=1), 1.0)
    # a profile constant on each DNS cell: the LES value is its exact length-weighted mean
    f = np.zeros(66)
    f[1:65] = np.sin(3 * dns["gyf"][1:65])
    got = yf.gyf @ f
    for j in (2, 10, 31):
        lo, hi = les["gy"][j], les["gy"][j + 1]
        acc = 0.0
        for k in range(1, 65):
            a, b = max(lo, dns["gy"][k]), min(hi, dns["gy"][k + 1])
            if b > a:
                acc += (b - a) * f[k]
        assert abs(got[j] - acc / (hi - lo)) < 1e-13


def test_stress_of_a_single_mode_and_of_a_constant():
    dcfg = configs.front_les(64, 33, 32, 5000.0, 1e-3, 0.1).copy(use_les=False).derive()
    lcfg = configs.front_les(16, 17, 16, 5000.0, 1e-3, 0.1).copy(les_model_type=6).derive()
    pd, pl = Problem(dcfg), Problem(lcfg)
    shape = (dcfg.ny + 2, dcfg.nz, dcfg.nx)
    x = np.arange(dcfg.nx) * dcfg.lx / dcfg.nx
    k = 2 * np.pi * 3 / dcfg.lx
    u1 = np.broadcast_to(np.cos(k * x), shape).copy()
    zero = np.zeros(shape)
    one = np.ones(shape)
    f = sgs_data.filter_snapshot(dcfg, pd.met, {"u1": u1, "u2": zero, "u3": zero, "th": one}, lcfg, pl.met)
    assert np.all(f.kappa_star == 0)
    # reconstruct tau_11 the way filter_snapshot does, against the closed form
    fx = dcfg.nx // lcfg.nx
    yf = sgs_data.y_filters(pd.met, pl.met)
    F = lambda a: sgs_data.apply_y(yf.gyf, sgs_data.tophat_xz(a, fx, fx))
    tau = F(u1 * u1) - F(u1) ** 2
    xl = x[::fx]
    dx = dcfg.lx / dcfg.nx
    want = 0.5 + 0.5 * H(2 * k, dx, fx) * np.cos(2 * k * xl) - H(k, dx, fx) ** 2 * np.cos(k * xl) ** 2
    assert np.allclose(tau[5, 3], want, atol=1e-12)
    c = sgs_data.filter_snapshot(dcfg, pd.met, {"u1": one, "u2": zero, "u3": 2 * one, "th": one}, lcfg, pl.met)
    assert np.all(c.nu_star == 0) and np.all(c.kappa_star == 0)
