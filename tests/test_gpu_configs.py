"""The workload configurations at their specified sizes and step counts.

- Taylor-Green 32^3, Newtonian, 200 steps: bit-exact against the oracle,
  with the kinetic-energy decay series.
- Kolmogorov FENE-P 64^3: FP64 bit-exact against the oracle over 20 steps;
  the mixed form over 1000 steps against the FP64 GPU step (periodic grid
  with nx a multiple of 64: the z-streamed mixed kernels with periodic y).
- Channel with one cylinder 128x64x64: FFT + line solve bit-exact over 10
  steps, SOR at its solver tolerance over 10 steps, 2000-step runs of the
  FP64 and mixed forms of both Poisson variants checked by the flow's
  invariants, and the mixed forms against FP64 and the FP32-storage floor.
"""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from fortran8_amd import api, configs
from fortran8_amd.fields import interior, interior_c, rel_l2
from fortran8_amd.params import PREC_MIXED, POISSON_SOR
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def assert_bitwise(o, g, tol=0.0):
    """Every field and history bit-identical (tol = 0), or within tol relative
    L2 on a grid with walls: next to a wall the C ghost is the linear
    extrapolation 2 C1 - C2, so the three candidate slopes of the C+-
    selection are equal up to rounding and the reference's own choice among
    them is decided by the last bit of acos and cos, in which the device's
    functions and the C library's differ.  The face values chosen then differ
    by a few ulps, and so do C and what follows from it."""
    so, sg = o.get_state(), g.get_state()
    pairs = [(n, interior(getattr(sg, n)), interior(getattr(so, n))) for n in ("U", "V", "W", "P")]
    pairs += [("C", interior_c(sg.C), interior_c(so.C))]
    pairs += [("hist%d" % q, sg.hist[q], so.hist[q]) for q in range(9)]
    for name, a, b in pairs:
        if tol:
            assert rel_l2(a, b) <= tol, (name, rel_l2(a, b))
        else:
            assert np.array_equal(a, b), (name, rel_l2(a, b))
    if tol:
        print("relative L2:", {name: "%.1e" % rel_l2(a, b) for name, a, b in pairs[:5]}, flush=True)


def storage_floor_run(config, steps):
    """Mixed vs FP64 and the FP32 floors (tests/_storage_floor.py)."""
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_storage_floor.py")
    r = subprocess.run([sys.executable, "-u", script, config, str(steps)], capture_output=True, text=True,
                       timeout=280)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return json.loads(r.stdout.strip().splitlines()[-1])


def report(res):
    keys = ("U", "V", "W", "P", "C", "u", "V_u", "W_u", "P_dyn", "drag", "ke")
    for run in ("mixed", "floor_ic", "floor_storage"):
        print("%-14s" % run + " ".join("%s=%.2e" % (k, res[run][k]) for k in keys if k in res[run]), flush=True)


def field_errors(ref, got):
    """Relative L2 of each field and of the stacked velocity."""
    e = {n: rel_l2(interior(getattr(got, n)), interior(getattr(ref, n))) for n in ("U", "V", "W", "P")}
    e["C"] = rel_l2(interior_c(got.C), interior_c(ref.C))
    e["u"] = rel_l2(np.concatenate([interior(getattr(got, n)).ravel() for n in "UVW"]),
                    np.concatenate([interior(getattr(ref, n)).ravel() for n in "UVW"]))
    return e


def test_cfg1_taylor_green_32_200_steps_bitwise_with_energy_decay():
    case = configs.taylor_green(n=32)
    print("cfg1 taylor-green 32^3, 200 steps", flush=True)
    o, g = po.load_case(case), api.load_case(case)
    keo, keg = [o.diagnostics()["ke"]], [g.diagnostics()["ke"]]
    for _ in range(10):
        o.step(20)
        g.step(20)
        keo.append(o.diagnostics()["ke"])
        keg.append(g.diagnostics()["ke"])
    assert_bitwise(o, g)
    # the same states; the reductions sum in different orders
    np.testing.assert_allclose(keg, keo, rtol=1e-12)
    ke = np.array(keg)
    assert np.all(np.diff(ke) < 0)
    # Viscous decay of the |k|^2 = 3 mode: K(t) = K(0) exp(-2 nu k_h^2 t), with
    # the 7-point Laplacian's eigenvalue k_h^2 = 3 (2 - 2 cos h)/h^2; the
    # nonlinear transfer out of the mode adds < 0.5 % by t = 1 at Re = 100.
    h = 2 * math.pi / 32
    kh2 = 3 * (2 - 2 * math.cos(h)) / h ** 2
    t = np.arange(11) * 20 * 0.005
    model = ke[0] * np.exp(-2 * kh2 * t / 100.0)
    print("KE series", ke.tolist(), "viscous model", model.tolist(), flush=True)
    assert np.max(np.abs(ke / model - 1)) < 5e-3


def test_cfg2_kolmogorov_64_fp64_twenty_steps_bitwise():
    case = configs.kolmogorov(n=64)
    print("cfg2 kolmogorov 64^3 FP64, 20 steps", flush=True)
    o, g = po.load_case(case), api.load_case(case)
    o.step(20)
    g.step(20)
    assert_bitwise(o, g)


@pytest.mark.timeout(300)
def test_cfg2_kolmogorov_64_mixed_1000_steps_against_fp64():
    """The z-streamed mixed kernels on a periodic grid (nx a multiple of 64).
    The flow is the laminar Kolmogorov profile in U with small V, W and P,
    so those three are compared on the velocity and dynamic-pressure scales.
    Every field is rounded to FP32 once per step (the increment form of the
    predictor, the Helmholtz solve, the correction and CStar/Newton), so the
    mixed step sits on the FP32-storage floor of this laminar flow: measured
    0.96-1.0 times it in every quantity (before: C 18 times, from the FP32
    sum of Cstar's 1/6 weights)."""
    print("cfg2 kolmogorov 64^3 mixed vs FP64, 1000 steps", flush=True)
    res = storage_floor_run("kolmogorov64", 1000)
    report(res)
    m, fl = res["mixed"], res["floor_storage"]
    assert m["sel_differ"] == 0
    for k in ("U", "C", "u", "V_u", "W_u", "P_dyn", "ke"):
        assert m[k] < 1e-4, k
        assert m[k] <= 1.5 * fl[k], (k, m[k], fl[k])


def test_cfg3_channel_cylinder_fft_line_ten_steps_bitwise():
    case = configs.channel_cylinder()
    print("cfg3 channel 128x64x64 FFT + line solve, 10 steps", flush=True)
    o, g = po.load_case(case), api.load_case(case)
    o.step(10)
    g.step(10)
    assert_bitwise(o, g, tol=1e-14)
    for a, b in zip(o.marker_forces(), g.marker_forces()):
        assert rel_l2(a, b) <= 1e-12


@pytest.mark.timeout(400)
def test_cfg3_channel_cylinder_sor_ten_steps_at_solver_tolerance():
    case = configs.channel_cylinder(poisson_variant=POISSON_SOR)
    case.params.sor_eps = 1e-9
    case.params.sor_itrmax = 50000
    print("cfg3 channel 128x64x64 SOR (eps 1e-9), 10 steps", flush=True)
    o, g = po.load_case(case), api.load_case(case)
    for s in range(10):
        o.step(1)
        g.step(1)
        print("  step", s + 1, "sor iterations oracle", o.diagnostics()["sor_itr"], "gpu",
              g.diagnostics()["sor_itr"], flush=True)
    so, sg = o.get_state(), g.get_state()
    e = field_errors(so, sg)
    # with Neumann walls Phi, and so P, is defined up to a constant, which the
    # two sweep orders leave at different values: compare P about its mean
    po_, pg = interior(so.P), interior(sg.P)
    e["P_zero_mean"] = rel_l2(pg - pg.mean(), po_ - po_.mean())
    print(json.dumps({k: float(v) for k, v in e.items()}), flush=True)
    for k in ("U", "V", "W", "C", "u"):
        assert e[k] < 1e-7, k
    assert e["P_zero_mean"] < 1e-6


def channel_invariants(g, label, steps, every):
    """div u at round-off (FFT) or solver tolerance (SOR), no NaN, tr C
    within (0, L^2), finite drag, CFL below the CStar convexity bound."""
    lp2 = g.p.lp ** 2
    for s in range(every, steps + 1, every):
        g.step(every)
        d = g.diagnostics()
        print(label, "step", s, {k: d[k] for k in ("ke", "drag", "div_max", "trc_min", "trc_max", "cfl")},
              flush=True)
        assert d["nan_flag"] == 0
        assert 0.0 < d["trc_min"] and d["trc_max"] < lp2
        assert math.isfinite(d["drag"]) and d["drag"] > 0.0
        assert d["cfl"] < 1.0
    return d


def test_cfg3_fp64_fft_line_2000_steps_invariants():
    g = api.load_case(configs.channel_cylinder())
    d = channel_invariants(g, "cfg3 FP64 FFT", 2000, 500)
    assert d["div_max"] < 1e-9


@pytest.mark.timeout(300)
def test_cfg3_fp64_sor_2000_steps_invariants():
    g = api.load_case(configs.channel_cylinder(poisson_variant=POISSON_SOR))
    d = channel_invariants(g, "cfg3 FP64 SOR", 2000, 500)
    assert 0 < d["sor_itr"] < g.p.sor_itrmax
    assert d["div_max"] < 1e-3


def test_cfg3_mixed_fft_line_2000_steps_invariants():
    mc = configs.channel_cylinder()
    mc.params.precision = PREC_MIXED
    g = api.load_case(mc)
    d = channel_invariants(g, "cfg3 mixed FFT", 2000, 500)
    assert d["div_max"] < 1e-3


def test_cfg3_mixed_sor_twenty_steps_against_fp64_sor():
    """The FP32 sweeps on the FP64 defect take the reference's stopping test:
    per step the same sweep count as the FP64 SOR (to a sweep or two where the
    test is crossed within rounding), and the same fields to FP32 rounding."""
    ref = api.load_case(configs.channel_cylinder(poisson_variant=POISSON_SOR))
    mc = configs.channel_cylinder(poisson_variant=POISSON_SOR)
    mc.params.precision = PREC_MIXED
    g = api.load_case(mc)
    for s in range(20):
        ref.step(1)
        g.step(1)
        i_ref, i_mix = ref.diagnostics()["sor_itr"], g.diagnostics()["sor_itr"]
        print("step", s + 1, "sweeps FP64", i_ref, "mixed", i_mix, flush=True)
        assert i_mix > 0
        assert abs(i_mix - i_ref) <= 0.05 * i_ref + 2
    sr, sg = ref.get_state(), g.get_state()
    e = field_errors(sr, sg)
    # W is the noise of the initial state alone: against the velocity scale
    un = np.linalg.norm(np.concatenate([interior(getattr(sr, n)).ravel() for n in "UVW"]))
    e["W_u"] = np.linalg.norm(interior(sg.W) - interior(sr.W)) / un
    print({k: "%.1e" % v for k, v in e.items()}, flush=True)
    for k in ("U", "V", "C", "u", "W_u"):
        assert e[k] < 1e-5, k
    d = g.diagnostics()
    assert d["nan_flag"] == 0 and d["div_max"] < 1e-3


@pytest.mark.timeout(300)
def test_cfg3_mixed_sor_2000_steps_invariants():
    mc = configs.channel_cylinder(poisson_variant=POISSON_SOR)
    mc.params.precision = PREC_MIXED
    g = api.load_case(mc)
    d = channel_invariants(g, "cfg3 mixed SOR", 2000, 500)
    assert 0 < d["sor_itr"] < g.p.sor_itrmax
    assert d["div_max"] < 1e-3


@pytest.mark.timeout(300)
@pytest.mark.parametrize("variant", ["channel", "channel_sor"])
def test_cfg3_mixed_2000_steps_against_fp64_and_the_storage_floor(variant):
    print("cfg3 channel 128x64x64 mixed vs FP64, 2000 steps,", variant, flush=True)
    res = storage_floor_run(variant, 2000)
    report(res)
    m, fl = res["mixed"], res["floor_storage"]
    assert m["sel_differ"] == 0
    for k in ("u", "P_dyn", "drag", "ke"):
        assert m[k] < 1e-4, k
    # The cylinder's wake is unstable at Re_D = 100 and amplifies any
    # perturbation: the floor is one sample of the FP32 rounding of the state,
    # the mixed step another, and their distances from FP64 differ by up to
    # about 2 from one sample to the next (measured 0.67-1.88 over the two
    # variants and two builds; the drag of the FP64 step started from the
    # FP32-rounded state alone moves by 4.4e-5).
    for k in ("U", "V", "W", "P", "C"):
        print("%s: mixed / FP32-storage floor = %.2f" % (k, m[k] / fl[k]), flush=True)
        assert m[k] <= 2.5 * fl[k], k
