"""Host helpers of diablo3_amd.config_runs on the oracle (CPU): the divergence
measure vanishes on a projected state and not on the raw initial condition; the
mean buoyancy profile is the kx = kz = 0 mode; compare() is zero on equal states."""
import numpy as np

import oracle
from diablo3_amd import config_runs, configs
from diablo3_amd.problem import Problem


def test_divergence_and_mean_profile():
    p = Problem(configs.small_dns(16, 17, 16))
    s = p.initial_state(seed=3, theta_noise=1e-3)
    assert config_runs.divergence(p, s) > 1e-6
    o = oracle.Oracle(p)
    o.upload_state(s)
    o.rem_div()
    sp = o.download_state()
    assert config_runs.divergence(p, sp) < 1e-13
    m = config_runs.mean_theta(p, sp)
    assert m.shape == (p.cfg.ny,)
    assert np.allclose(m, sp.cth[0][1:p.cfg.ny + 1, 0, 0].real)
    c = config_runs.compare(p, sp, p, sp)
    assert all(v == 0.0 for v in c.values())
