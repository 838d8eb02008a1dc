"""Discrete adjoint of the FP64 step (fp_adj_record / fp_adj_backward and
the torch.autograd.Function in fortran8_amd.adjoint) against central finite
differences of the same step with the recorded C+- selections."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

SCRIPT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_adjoint_check.py")


def run_check(which, timeout=600):
    print("adjoint check", which, flush=True)
    r = subprocess.run([sys.executable, SCRIPT, which], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)


def test_adjoint_one_step_vjp_matches_finite_differences():
    out = run_check("vjp")
    assert out["rel_h1e-05"] < 1e-8
    assert out["rel_uvwp"] < 1e-6 and out["rel_c"] < 1e-6 and out["rel_hist"] < 1e-6


def test_adjoint_parameter_and_force_gradients_multi_segment():
    # stretched conformation (tr C ~ 30 of L^2 = 100), Newton converged to
    # round-off; reference derivative: Richardson-extrapolated central
    # differences (h = 1e-4, 5e-5)
    out = run_check("params")
    for k in ("rel_wi", "rel_beta", "rel_L2", "rel_F"):
        assert out[k] < 1e-6, k


def test_adjoint_cfg5_drag_gradients_and_cost():
    out = run_check("cfg5", timeout=900)
    # configuration 5 as specified (reference Newton tolerance) and with the
    # Newton converged to round-off; reference derivative: Richardson-
    # extrapolated central differences.  dJ/d(L^2) vanishes here (tr C ~ 3):
    # it is checked on the stretched state of the multi-segment test.
    for tag in ("ref", "tight"):
        assert out[tag + "_rel_wi_rich"] < 1e-5, tag
        assert out[tag + "_rel_beta_rich"] < 1e-6, tag
        assert out[tag + "_rel_F"] < 1e-6, tag
    assert out["ratio_backward_forward"] < 3.0
