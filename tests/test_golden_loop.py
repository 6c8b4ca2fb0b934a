"""The full-size loop fixture (tests/golden/loop_oracle.json) belongs to the
problem the generator makes today, and its one-ulp sensitivity shows the
loop's chaotic step: the twins agree to 1e-8 for the first 8 iterations and
separate by more than 1e-7 in the next one."""
import json
import os

import numpy as np

from golden.make_loop_fixture import problem_digest
from nav24_amd import problems

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "loop_oracle.json")


def test_fixture_is_for_todays_problem():
    with open(PATH) as f:
        fx = json.load(f)
    p = problems.workload(2, seed=0)
    assert (p.n_cam, p.n_pt, p.n_obs) == (fx["n_cam"], fx["n_pt"], fx["n_obs"])
    assert problem_digest(p) == fx["problem_sha1"]
    assert len(fx["cost_history"]) == fx["iterations"] + 1
    assert np.asarray(fx["T_cw"]).shape == (p.n_cam, 12)
    sens = np.asarray(fx["one_ulp_sensitivity"]["history"])
    assert np.all(sens[:8] <= 1e-8) and sens[8] > 1e-7
