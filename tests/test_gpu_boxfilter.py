"""dgpu_box_filter (the learned closure's training filter on the device) against
the numpy restatement sgs_data.filter_fields on the same state: every one of the
14 filtered fields within 1e-13 of the largest value of that field, for LES grids
that nest (y rows merged 2:1) and that do not, and x/z filter ratios 2, 4 and 8."""
import numpy as np
import pytest

import oracle
from diablo3_amd import configs, sgs_data
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dns,les", [((64, 33, 64), (16, 17, 16)), ((64, 33, 32), (32, 13, 4)),
                                     ((32, 65, 64), (16, 33, 8))])
def test_box_filter_matches_numpy(dns, les):
    d = configs.front_les(*dns, 5000.0, 1e-3, 0.1, n2=1.0).copy(use_les=False)
    l = configs.front_les(*les, 5000.0, 1e-3, 0.1, n2=1.0)
    pd, pl = Problem(d), Problem(l)
    s = pd.initial_state(seed=3, theta_noise=0.05)
    rng = np.random.default_rng(5)
    for a in (s.cu1, s.cu2, s.cu3, s.cth[0]):
        a[1:-1] += 0.05 * (rng.standard_normal(a[1:-1].shape) + 1j * rng.standard_normal(a[1:-1].shape))
    o = oracle.Oracle(pd)
    phys = {"u1": o.fft_to_physical(s.cu1), "u2": o.fft_to_physical(s.cu2),
            "u3": o.fft_to_physical(s.cu3), "th": o.fft_to_physical(s.cth[0])}
    want = sgs_data.filter_fields(d, pd.met, phys, l, pl.met)
    g = Solver(pd)
    g.upload_state(s)
    mf, mg = sgs_data.y_weights(pd.met, pl.met)
    got = g.box_filter(d.nx // l.nx, d.nz // l.nz, mf, mg)
    assert got.shape == want.shape
    errs = [float(np.abs(got[f] - want[f]).max() / max(np.abs(want[f]).max(), 1e-300)) for f in range(14)]
    print(dns, les, "worst field", sgs_data.BOX_FIELDS[int(np.argmax(errs))], f"{max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-13
    # the state is not changed
    sg = g.download_state()
    assert np.array_equal(sg.cu1, s.cu1) or np.abs(sg.cu1 - s.cu1).max() == 0.0
    g.close()
