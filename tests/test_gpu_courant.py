"""variable_dt: courant (for/flow.f90:746-857) evaluated by the device inside
rk_step 3 (for/solver.f90:1241-1244) against the oracle, through the C ABI.

Bar: delta_t within 1e-13 relative of the oracle's (the CFL limit is a min over
points of quotients of the velocity, which the two sides hold to round-off);
the event scheduling and the update_dt cadence identical.
"""
import numpy as np
import pytest

import oracle
from diablo3_amd import configs
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("les", [False, True])
def test_courant_matches_oracle(les):
    if les:
        cfg = configs.front_les(32, 33, 32, 5000.0, 2e-4, 1e-2)
    else:
        cfg = configs.small_dns(32, 33, 32)
    p = Problem(cfg)
    o = oracle.Oracle(p, threads=16)
    s = p.initial_state(seed=5, theta_noise=1e-3)
    rng = np.random.default_rng(5)
    sh = (cfg.ny - 3, 3, 3)
    for a in (s.cu1, s.cu2, s.cu3):
        a[2:cfg.ny - 1, 1:4, 1:4] += 0.3 * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    o.upload_state(s)
    o.rem_div()
    s0 = o.download_state()
    g = Solver(p)
    g.upload_state(s0)
    update_dt, cfl = 2, 0.5
    g.set_variable_dt(True, update_dt, cfl)
    o.set_variable_dt(True, update_dt, cfl)
    t, dt = 0.0, cfg.delta_t
    # time_limit, save_stats_time, save_flow_time; the event times move on as
    # the main loop moves them after a save (for/diablo.f90:87-121)
    time_limit, stats_dt, flow_dt = 10.0, 0.05, 1.0
    stats_t, flow_t = stats_dt, flow_dt
    cnf = not les
    for n in range(1, 5):
        for rk in (1, 2, 3):
            g.rk_substep(rk, delta_t=dt, time=t, time_step=n, create_new_flow=cnf)
            o.rk_substep(rk, delta_t=dt, time=t, time_step=n, create_new_flow=int(cnf))
        events = (time_limit, stats_t, flow_t)
        dg, rg_ = g.courant(t, *events)
        do, ro_ = o.courant(t, *events)
        print(f"step {n}: ran {rg_}/{ro_} dt {dg!r} {do!r}", flush=True)
        assert rg_ == ro_ == (n % update_dt == 0)
        if ro_:
            assert do > 0 and abs(dg - do) <= 1e-13 * do, (dg, do)
            assert dg != dt  # the CFL limit or an event binds for this state
            dt = do
        t += dt
        if t >= stats_t:
            stats_t += stats_dt
        if t >= flow_t:
            flow_t += flow_dt
    a, b = g.download_state(), o.download_state()
    assert np.linalg.norm(a.cu1 - b.cu1) <= 1e-11 * np.linalg.norm(b.cu1)
