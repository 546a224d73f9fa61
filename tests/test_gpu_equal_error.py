"""bench.py --equal-error on a small state (36^3 extended, 3 AC + 1 CH, 8
nuclei): the plain FP64 path is first order in dt against its Richardson
reference, and the headline path's error and the matching FP64 step come out
finite, so the study behind the 192^3 figure keeps working."""
import os
import sys
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import pyetd  # noqa: E402
from pyetd import controller as CT  # noqa: E402
from pyetd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def test_equal_error_study_on_a_small_state():
    print("equal-error study, n = 36", flush=True)
    prm = W.scaled(W.beta_alpha_192(), 36, nac=3, eps0=W.eigenstrains(3, 0.05, 5))
    args = types.SimpleNamespace(weights=bench.CONTROLLER, controller="learned", tol=1e-4, precision=2)
    d = bench.equal_error(args, pyetd, W, CT, prm, nsteps=20, levels=5)
    rows = d["fixed_dt_fp64"]
    assert [r["steps"] for r in rows] == [20, 40, 80, 160, 320]
    errs = [r["max_deta"] for r in rows]
    ratios = [a / b for a, b in zip(errs[:3], errs[1:4])]  # rows outside the reference
    print(f"  errors {errs}, ratios {ratios}", flush=True)
    assert all(1.8 < q < 2.2 for q in ratios)  # first order
    h = d["headline"]
    assert h["energy_monotone"] and np.isfinite(h["max_deta"]) and h["max_deta"] > 0
    assert np.isfinite(d["fp64_dt_at_headline_error"]) and d["fp64_dt_at_headline_error"] > 0
    assert np.isfinite(d["speedup_at_equal_error"]) and d["speedup_at_equal_error"] > 0
