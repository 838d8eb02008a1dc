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


def test_adjoint_one_step_vjp_with_smac_spreading_weight():
    out = run_check("vjp_smac_weight")
    assert out["params"]["ibm_weight"] > 0
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
    # configuration 5 with the Newton converged (8 iterations, round-off):
    # central differences at h = 1e-6 within 1e-6 of the adjoint.  dJ/d(L^2)
    # is about 4e-14 here (tr C ~ 3) and is checked on the stretched state of
    # the multi-segment test.
    for k in ("conv_rel_wi_h1e-06", "conv_rel_beta_h1e-06", "conv_rel_F"):
        assert out[k] < 1e-6, k
    # and the complex-step gradient of the FP64 CPU oracle (the same discrete
    # branch: the oracle and the GPU step are bit-identical on this periodic
    # case), at both Newton settings and for all four directions, L^2
    # included (dJ/dL^2 ~ 1e-9, below what differences resolve): measured
    # 4e-16 .. 7e-14
    for tag in ("conv", "ref"):
        for k in ("wi", "beta", "L2", "F"):
            key = "%s_rel_%s_cs" % (tag, k)
            print(key, out[key])
            assert out[key] < 1e-10, key
    # the reference's tolerance (1e-6) stops the Newton at a point that
    # moves in steps with the parameters: Richardson-extrapolated differences
    # at h = 1e-4 within 1e-5
    assert out["ref_rel_wi_rich"] < 1e-5 and out["ref_rel_beta_rich"] < 1e-6 and out["ref_rel_F"] < 1e-6
    # the reverse sweep alone (recomputation included) costs below 3 forward
    # runs; recording plus reverse is reported beside it
    assert out["ref_ratio_backward_forward"] < 3.0
    print("record + reverse / forward:", out["ref_ratio_record_backward_forward"])
    # mixed-precision forward: the gradients agree with the FP64 ones to the
    # accuracy of the mixed trajectory
    for k in ("mixed_rel_wi_vs_fp64", "mixed_rel_beta_vs_fp64", "mixed_rel_F_vs_fp64"):
        print(k, out[k])
        assert out[k] < 1e-2, k


def test_adjoint_taped_last_segment_equals_its_recomputation():
    """The FP64 recording tapes its last checkpoint segment, which the reverse
    sweep then takes instead of recomputing it; a setter call after the
    recording (here the physics, set to the same values) drops that tape, so
    the sweep recomputes every segment.  Both give the same gradients."""
    import numpy as np
    from fortran8_amd import api, configs
    case = configs.adjoint_case()
    p = case.params
    nsteps, every = 60, 20
    w = np.zeros(nsteps)
    w[nsteps // 2:] = 1.0 / (nsteps - nsteps // 2)
    g = api.load_case(case)
    out = []
    for drop in (False, True):
        g.set_state(case.state)
        g.adj_record(nsteps, every)
        if drop:
            g.set_physics(p.wi, p.beta, p.lp)
        out.append(g.adj_backward(weights=w))
    g.close()
    print("taped", out[0], "recomputed", out[1], flush=True)
    assert np.all(np.isfinite(out[0])) and np.any(out[0] != 0.0)
    assert np.allclose(out[0], out[1], rtol=1e-12, atol=0.0)
