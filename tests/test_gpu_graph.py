"""dgpu_step as a captured HIP graph against the same step run as direct launches.

Two handles from one seeded state, one with the graph (the default) and one
without, stepped side by side: every field bit-identical after every step, with
eager device work (statistics, downloads) of the program between the replays; a
change of delta_t and the LES spin-up gate (time_step 100 -> 101 of a new flow)
recapture the graph; an LES handle and a mixed-precision handle replay too.
"""
import numpy as np
import pytest

from diablo3_amd import configs
from diablo3_amd.config import MIXED
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver

from test_gpu_les import fields, les_problem

pytestmark = pytest.mark.gpu


def _pair(p, s0):
    a, b = Solver(p), Solver(p)
    b.set_graph(False)
    a.upload_state(s0)
    b.upload_state(s0)
    return a, b


def _same(a, b, what):
    fa, fb = fields(a.download_state()), fields(b.download_state())
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), (what, k, float(np.abs(fa[k] - fb[k]).max()))


def _run(a, b, schedule, cnf):
    t = 0.0
    for n, (dt, ts) in enumerate(schedule):
        a.step(delta_t=dt, time=t, time_step=ts, create_new_flow=cnf)
        b.step(delta_t=dt, time=t, time_step=ts, create_new_flow=cnf)
        _same(a, b, f"step {n} (dt {dt}, time_step {ts})")
        # eager device work between the replays (save_stats_chan also applies the
        # BC posts to the state, so both handles take it)
        a.stats(time=t)
        b.stats(time=t)
        t += dt


@pytest.mark.parametrize("precision", [0, MIXED])
def test_graph_dns_matches_direct_launches(precision):
    cfg = configs.small_dns(32, 33, 32, precision=precision)
    p = Problem(cfg)
    a = Solver(p)
    s0 = p.initial_state(seed=2, theta_noise=1e-3)
    a.upload_state(s0)
    a.rem_div()
    s0 = a.download_state()
    a.close()
    a, b = _pair(p, s0)
    dt = cfg.delta_t
    _run(a, b, [(dt, 1), (dt, 2), (dt, 3), (0.5 * dt, 4), (0.5 * dt, 5), (dt, 6)], cnf=True)


def test_graph_les_gate_and_replays():
    p, s0 = les_problem(32, 33, 32)
    a, b = _pair(p, s0)
    dt = p.cfg.delta_t
    # new flow: DNS through time_step 100, LES from 101 on (for/solver.f90:615)
    _run(a, b, [(dt, 99), (dt, 100), (dt, 101), (dt, 102), (dt, 103)], cnf=True)
    nt_a, kt_a = a.download_les()
    nt_b, kt_b = b.download_les()
    assert nt_a.max() > 0 and np.array_equal(nt_a, nt_b) and np.array_equal(kt_a[0], kt_b[0])
