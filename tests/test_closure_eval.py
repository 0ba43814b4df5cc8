"""The a-posteriori metrics of diablo3_amd/closure_eval.py on constructed states:
zero error for identical states; the mean buoyancy profile is the (0, 0) mode of
theta; a uniform scaling of the velocity by a shifts the TKE spectrum by
log10(a^2) at every kx."""
import numpy as np

from diablo3_amd import closure_eval, configs
from diablo3_amd.problem import Problem


def state(seed):
    cfg = configs.front_les(16, 17, 16, 5000.0, 1e-3, 0.1)
    p = Problem(cfg)
    s = p.initial_state(seed=seed, theta_noise=1e-2)
    rng = np.random.default_rng(seed)
    for a in (s.cu1, s.cu2, s.cu3):
        a += 0.1 * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape))
    return cfg, s


def test_identical_states_have_zero_error():
    cfg, s = state(1)
    e = closure_eval.compare(s, s, cfg)
    assert e["theta_profile_rel_l2"] == 0.0 and e["tke_spectrum_rms_log10"] == 0.0


def test_profile_and_spectrum_closed_forms():
    cfg, s = state(2)
    assert np.array_equal(closure_eval.mean_profile(s, cfg), s.cth[0][1:cfg.ny, 0, 0].real)
    b = s.copy()
    for name in ("cu1", "cu2", "cu3"):
        setattr(b, name, 3.0 * getattr(s, name))
    e = closure_eval.compare(b, s, cfg)
    assert abs(e["tke_spectrum_rms_log10"] - np.log10(9.0)) < 1e-12
    assert e["theta_profile_rel_l2"] == 0.0
