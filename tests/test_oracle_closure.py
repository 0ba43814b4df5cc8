"""Pins for the oracle's restatement of the learned closure's per-point stage
(oracle/closure.py) and its type-6 hooks (oracle_les_capture, oracle_les_override).

* features of tensors whose invariants are known in closed form: pure strain,
  pure rotation, axisymmetric strain, a scalar gradient along an eigenvector;
* the override with every point on the fallback (the type-5 coefficient with its
  sign bit set) steps exactly as type 5; with a constant positive coefficient
  nu_t = -(1/2) delta(j) c on the rows ghost_les_mpi keeps (for/les.f90:2229-2246).
"""
import numpy as np

import oracle
from oracle import closure as ocl
from diablo3_amd.problem import Problem

from test_oracle_les import les_cfg


def les_cfg_small():
    return les_cfg(16, 17, 16)


def les_state(p, amp=0.05, seed=11):
    """A projected front state with resolved low modes large enough that the
    closure is active on most points."""
    cfg = p.cfg
    s = p.initial_state(seed=seed, theta_noise=1e-3)
    rng = np.random.default_rng(seed)
    shp = (cfg.ny - 3, 4, 4)
    for a in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        a[2:cfg.ny - 1, 1:5, 1:5] += amp * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    o = oracle.Oracle(p, threads=4)
    o.upload_state(s)
    o.rem_div()
    return o.download_state()


class _Cfg:
    nx = nz = 8
    ny, ny2 = 3, 5
    lx = lz = 8.0
    grav_y = 1.0
    nu = 1e-3


def _one_point(a=(0, 0, 0, 0, 0, 0), o=(0, 0, 0), b=(0, 0, 0), ty=0.0):
    cap = np.zeros((5, 1, 1, 16))
    cap[2, 0, 0, 0:6] = a
    cap[2, 0, 0, 6:9] = o
    cap[2, 0, 0, 10:13] = b
    cap[2, 0, 0, 13] = ty
    dy = np.array([0.0, 0.1, 0.2, 0.2, 0.1])
    X, s2 = ocl.features(cap, _Cfg, dy)
    return X[2, 0, 0].astype(np.float64), float(s2[2, 0, 0])


def test_features_closed_forms():
    # pure strain diag(2, -1, -1): s2 = 6, no rotation -> f0 = -1, f1 = tr S^3 / s2^1.5 = 6 / 6^1.5
    f, s2 = _one_point(a=(2, -1, -1, 0, 0, 0))
    assert s2 == 6.0 and f[0] == -1.0
    assert abs(f[1] - 6.0 / 6.0 ** 1.5) < 1e-6 and f[2] == 0 and f[4] == 0
    assert abs(f[3] - (-6.0 / (6.0 ** 0.5 * 6.0))) < 1e-6      # -(tr S^3)/(s2^.5 e)
    # pure rotation: s2 = 0 -> every ratio with s2 below is 0, f0 = (w2 - 0)/w2 = 1
    f, s2 = _one_point(o=(0.5, 0.0, 0.0))
    assert s2 == 0.0 and f[0] == 1.0 and np.all(f[1:4] == 0)
    # strain diag(1, -1, 0) with rotation about z (o4): S P has trace -w2/2 * (1 - 1) = 0
    f, _ = _one_point(a=(1, -1, 0, 0, 0, 0), o=(0.5, 0, 0))
    assert abs(f[2]) < 1e-7 and abs(f[0] - (0.5 - 2.0) / 2.5) < 1e-7
    # scalar gradient along the strain eigenvector e_y (eigenvalue -1): b.S.b/(|b|^2 s2^.5) = -1/sqrt(2)
    f, _ = _one_point(a=(1, -1, 0, 0, 0, 0), b=(0, 2.0, 0), ty=5.0)
    assert f[5] == 1.0 and abs(f[6] + 1.0 / np.sqrt(2.0)) < 1e-6 and abs(f[7] - 0.5) < 1e-6
    assert abs(f[9] - np.tanh(5.0 / 2.0)) < 1e-6
    # filter-width features: log(1 + D2 sqrt(2 s2)/nu), log(delta_y^2/(delta_x delta_z)) at GY row 2
    dxs = 8.0 / 8 * 3.0
    d2 = 3.0 / (1 / dxs ** 2 + 1 / 0.4 ** 2 + 1 / dxs ** 2)
    assert abs(f[11] - np.log1p(d2 * 2.0 / 1e-3)) < 1e-5 and abs(f[12] - np.log(0.16 / dxs ** 2)) < 1e-6


def test_override_fallback_steps_as_type5_and_constant_coefficient():
    cfg = les_cfg_small()
    p = Problem(cfg)
    s0 = les_state(p)
    o5 = oracle.Oracle(p, threads=4)
    o5.upload_state(s0)
    cap = o5.les_capture()
    o5.rk_substep(1, time=0.0, time_step=1, create_new_flow=0)
    ref = o5.download_state()
    nt5, _ = o5.download_les()
    assert cap[2:cfg.ny + 1, ..., ocl.CAP_V].max() > 0
    # every point on the fallback
    ob = oracle.Oracle(p, threads=4)
    ob.upload_state(s0)
    ob.les_override(-cap[..., ocl.CAP_V], -cap[..., ocl.CAP_VT])
    ob.rk_substep(1, time=0.0, time_step=1, create_new_flow=0)
    got = ob.download_state()
    for a, b in ((got.cu1, ref.cu1), (got.cu2, ref.cu2), (got.cu3, ref.cu3), (got.cth[0], ref.cth[0]), (got.cp, ref.cp)):
        assert np.array_equal(a, b)
    assert np.array_equal(ob.download_les()[0], nt5)
    # a constant positive coefficient c everywhere
    c = 0.37
    oc = oracle.Oracle(p, threads=4)
    oc.upload_state(s0)
    full = np.full((cfg.ny2, cfg.nz, cfg.nx), c)
    oc.les_override(full, full)
    oc.rk_substep(1, time=0.0, time_step=1, create_new_flow=0)
    nt, kt = oc.download_les()
    dy = p.met["dy"]
    dxs, dzs = cfg.lx / cfg.nx * 3.0, cfg.lz / cfg.nz * 3.0
    j = np.arange(3, cfg.ny)
    delta = -2.0 * ocl.C_AMD ** 2 * 3.0 / (1 / dxs ** 2 + 1 / (2 * dy[j]) ** 2 + 1 / dzs ** 2)
    want = (-0.5 * delta * c)[:, None, None]
    assert np.abs(nt[j] - want).max() <= 1e-15 * want.max()
    assert np.abs(kt[0][j] - want).max() <= 1e-15 * want.max()
    assert np.all(nt[[1, 2, cfg.ny, cfg.ny + 1]] == 0)
