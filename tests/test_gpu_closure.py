"""les_model_type 6 (the learned closure) through the C ABI.

* With no network set, or with an empty training envelope, every point takes
  the AMD coefficient: the step must match the oracle's type 5 to the LES
  parity bar (1e-11 on every field, cp included).
* dgpu_les_features: the AMD column times the scale is the type-5 eddy
  viscosity at GYF, and the features are finite and bounded where defined.
* The in-envelope step against the oracle's network restatement, with
  dgpu_closure_counts against the oracle's envelope mask.
"""
import numpy as np
import pytest

import oracle
from diablo3_amd import closure, configs
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver, load

from test_gpu_les import fields, les_problem, rel, run_oracle

pytestmark = pytest.mark.gpu


def type6(p5):
    cfg = p5.cfg.copy(les_model_type=6)
    return Problem(cfg)


@pytest.mark.parametrize("envelope", ["none", "empty"])
def test_type6_fallback_matches_amd_oracle(envelope):
    p5, s0 = les_problem(32, 33, 32)
    p6 = type6(p5)
    g = Solver(p6)
    if envelope == "empty":
        norm = np.zeros(64, np.float32)
        norm[16:29] = 1.0
        norm[32:48] = 1.0   # lo > hi: no point is inside
        norm[48:64] = -1.0
        m = closure.Mlp(seed=1)
        g.set_closure(m.params, norm)
    g.upload_state(s0)
    t = 0.0
    for n in range(3):
        for rk in (1, 2, 3):
            g.rk_substep(rk, time=t, time_step=n + 1, create_new_flow=False)
        t += p5.cfg.delta_t
    o = run_oracle(p5, s0, 3)
    fg, fo = fields(g.download_state()), fields(o.download_state())
    fa = fields(run_oracle(p5, s0, 3, variant=1).download_state())
    cp_floor = rel(fa["cp"], fo["cp"])
    errs = {k: rel(fg[k], fo[k]) for k in fo}
    (ntg, ktg), (nto, kto) = g.download_les(), o.download_les()
    errs["nu_t"] = rel(ntg, nto)
    errs["kappa_t"] = rel(ktg[0], kto[0])
    print(envelope, {k: f"{v:.2e}" for k, v in errs.items()}, f"oracle cp self-floor {cp_floor:.1e}", flush=True)
    for k, e in errs.items():
        assert e <= 1e-11, (k, e)


def test_features_amd_column_is_type5_viscosity():
    p5, s0 = les_problem(32, 33, 32)
    g = Solver(type6(p5))
    g.upload_state(s0)
    X, sc, am, ins = g.les_features()
    cfg = p5.cfg
    n = (cfg.ny - cfg.jstart) * cfg.nz * cfg.nx
    assert X.shape == (n, 16) and np.all(ins == 1)
    assert np.all(np.isfinite(X)) and np.all(X[:, 13:] == 0)
    # bounded features: ratios in [-1, 1] (0, 5, 7, 8, 9, 10), others by construction
    for k in (0, 5, 9, 10):
        assert np.all(np.abs(X[:, k]) <= 1.0 + 1e-6), k
    for k in (4, 7, 8):
        assert np.all((X[:, k] >= -1e-6) & (X[:, k] <= 1.0 + 1e-6)), k
    # the AMD column: coef / sqrt(2 s2) = max(0, f3) / sqrt(2) on every point with s2 > 0
    live = sc > 0
    assert live.mean() > 0.9
    assert np.allclose(am[live, 0], np.maximum(0.0, X[live, 3]) / np.sqrt(2.0), rtol=1e-5, atol=1e-6)
    assert np.all(np.isfinite(am)) and np.all(am >= 0) and (am[:, 1] > 0).mean() > 0.05
    print("features: s2>0 on", f"{live.mean():.3f}", "AMD active on", f"{(am[:, 0] > 0).mean():.3f}", flush=True)


@pytest.mark.parametrize("rows_out", [False, True])
def test_type6_network_points_match_oracle(rows_out):
    """The in-envelope path: the device's features, network, envelope test and
    coefficient assembly (k_closure, k_amd<..., 6>) against the oracle's
    restatement (oracle/closure.py on the type-5 quantities the oracle captures,
    oracle/mlp.py's BF16-emulating forward pass, oracle_les_override).  The
    committed network with an envelope of the 1st-99th percentiles of this
    state's own features (each bound set in the gap between two neighbouring
    feature values, so that a last-bit difference of a feature cannot move a
    point across it), so that points both inside and outside are exercised;
    with rows_out the upper bound of the row-constant width-ratio feature is its
    median instead, so that whole rows leave the envelope (k_closure's waves with
    no point inside skip the network).
    Bars (BF16-aware: a feature that differs in its last float bit can round to a
    neighbouring BF16 input): the count of network points (dgpu_closure_counts)
    equal to the oracle's; nu_t and kappa_t after one substep within 2e-3
    (relative L2).  The
    state moves with the eddy viscosity linearly over one substep, so its error
    is held to 10x the linear response of the nu_t difference: the state change
    between the oracle's network and type-5 substeps, scaled by the ratio of the
    device-vs-oracle nu_t difference to the network-vs-type-5 one (and never
    tighter than the 1e-11 parity bar)."""
    from oracle import closure as ocl
    p5, s0 = les_problem(32, 33, 32)
    cfg = p5.cfg
    params, norm, src = closure.load_params()
    o = oracle.Oracle(p5, threads=16)
    o.upload_state(s0)
    cap = o.les_capture()
    o.rk_substep(1, time=0.0, time_step=1, create_new_flow=0)
    X, s2 = ocl.features(cap, cfg, p5.met["dy"])
    live = s2 > 0
    F = X[live][:, :ocl.mlp.NIN]
    env = np.asarray(norm, np.float32).copy()

    def gap_bound(v, frac):
        # a bound halfway between two neighbouring distinct values near the
        # quantile: a feature that differs in its last bits between two correct
        # implementations then lies on the same side of it
        u = np.unique(v)
        i = min(max(int(np.searchsorted(u, np.quantile(v, frac))), 1), len(u) - 1)
        return np.float32(0.5 * (float(u[i - 1]) + float(u[i])))

    for k in range(ocl.mlp.NIN):
        env[32 + k] = gap_bound(F[:, k], 0.01)
        env[48 + k] = gap_bound(F[:, k], 0.99)
    if rows_out:
        env[48 + 12] = gap_bound(F[:, 12], 0.5)
    cgy, kgy, inside = ocl.coefficients(cap, cfg, p5.met["dy"], params, env)
    frac = float(inside[live].mean())
    rows_in = int(inside.any(axis=(1, 2)).sum())
    print(f"network {src}: {live.sum()} GY points, inside the envelope {frac:.3f}, "
          f"rows with a point inside {rows_in} of {cfg.ny - 1}", flush=True)
    assert 0.2 < frac < 0.99
    if rows_out:
        assert rows_in < (cfg.ny - 1) * 0.7
    ob = oracle.Oracle(p5, threads=16)
    ob.upload_state(s0)
    ob.les_override(cgy, kgy)
    ob.rk_substep(1, time=0.0, time_step=1, create_new_flow=0)
    g = Solver(type6(p5))
    g.set_closure(params, env)
    g.upload_state(s0)
    g.rk_substep(1, time=0.0, time_step=1, create_new_flow=False)
    # dgpu_closure_counts: every GY point of rows 2..Nyp evaluated once, and the
    # network taken where the oracle's envelope test passes (up to the points a
    # last-bit feature difference moves across an envelope bound)
    nn, npts = g.closure_counts()
    n_in = int(inside.sum())
    print(f"closure counts: network {nn} of {npts}, oracle inside {n_in}", flush=True)
    assert npts == (cfg.ny - 1) * cfg.nz * cfg.nx
    assert nn == n_in
    (ntg, ktg), (nto, kto) = g.download_les(), ob.download_les()
    # which GY points took the network: nu_t of the device at GY over the
    # oracle's two candidates (network or fallback) on rows 3..Nyp-1
    j = slice(3, cfg.ny)
    fall = np.abs(cap[..., ocl.CAP_V])
    dfac = ntg[j] / np.where(nto[j] != 0, nto[j], 1.0)
    errs = {"nu_t": rel(ntg, nto), "kappa_t": rel(ktg[0], kto[0])}
    sg, so = fields(g.download_state()), fields(ob.download_state())
    errs.update({k: rel(sg[k], so[k]) for k in so})
    print({k: f"{v:.2e}" for k, v in errs.items()}, f"nu_t ratio median {np.median(dfac):.6f}", flush=True)
    # against the pure type-5 step the network changed the eddy viscosity
    d_nu = rel(nto, o.download_les()[0])
    assert d_nu > 1e-3 and fall[j].max() > 0
    assert errs["nu_t"] <= 2e-3 and errs["kappa_t"] <= 2e-3
    s5 = fields(o.download_state())
    for k in so:
        bar = max(1e-11, 10.0 * rel(so[k], s5[k]) * errs["nu_t"] / d_nu)
        print(f"{k}: device vs oracle {errs[k]:.2e}, network vs type 5 {rel(so[k], s5[k]):.2e}, "
              f"bar {bar:.2e}", flush=True)
        assert errs[k] <= bar, (k, errs[k], bar)
