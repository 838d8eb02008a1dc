"""Parity of the HIP step (through the C-ABI) with the FP64 CPU oracle.

The plain FP64 path evaluates every expression in the oracle's order without
fused multiply-add, so the bar is bit-exact equality of U, V, W, P and C
after several steps; the SOR variant (red-black on the GPU, lexicographic in
the reference) is compared at the solver tolerance."""
import math

import numpy as np
import pytest

from fortran8_amd import api, configs
from fortran8_amd.fields import interior, interior_c, rel_l2, cylinder_markers
from fortran8_amd.params import METHOD_EXPLICIT, METHOD_FFT, METHOD_IBM, POISSON_SOR, default_params
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "P")


def run_both(case, steps):
    print("parity case", case.name, case.params.nx, case.params.ny, case.params.nz, "steps", steps, flush=True)
    o = po.load_case(case)
    g = api.load_case(case)
    o.step(steps)
    g.step(steps)
    return o, g


def assert_bitwise(o, g):
    so, sg = o.get_state(), g.get_state()
    for name in FIELDS:
        a, b = interior(getattr(sg, name)), interior(getattr(so, name))
        assert np.array_equal(a, b), (name, rel_l2(a, b), np.max(np.abs(a - b)))
    a, b = interior_c(sg.C), interior_c(so.C)
    assert np.array_equal(a, b), ("C", rel_l2(a, b), np.max(np.abs(a - b)))
    assert np.array_equal(sg.hist, so.hist)


def assert_close(o, g, tol):
    so, sg = o.get_state(), g.get_state()
    for name in FIELDS[:3]:
        assert rel_l2(interior(getattr(sg, name)), interior(getattr(so, name))) < tol, name
    assert rel_l2(interior_c(sg.C), interior_c(so.C)) < tol


def test_taylor_green_newtonian_bitwise():
    o, g = run_both(configs.taylor_green(n=16), 10)
    assert_bitwise(o, g)


def test_kolmogorov_fenep_bitwise():
    o, g = run_both(configs.kolmogorov(n=32), 10)
    assert_bitwise(o, g)


def test_anisotropic_noncubic_grid_bitwise():
    case = configs.kolmogorov(n=32, nz=8)
    case.params.ny = 16
    case = configs.Case(case.name, case.params, _resize(case, 32, 16, 8), force=None)
    o, g = run_both(case, 5)
    assert_bitwise(o, g)


def _resize(case, nx, ny, nz):
    from fortran8_amd.configs import _spd_perturbed_c, _ghost_c, _noise_velocity
    from fortran8_amd.fields import State, pad_periodic, ghosted_zeros
    p = case.params
    p.nx, p.ny, p.nz = nx, ny, nz
    p.lx, p.ly, p.lz = 2.0, 1.0, 1.0
    u, v, w = _noise_velocity(9, nx, ny, nz, 1e-2, 2.0 / nx, 1.0 / ny, 1.0 / nz, False, 0)
    c = _spd_perturbed_c(9, nx, ny, nz, amp=0.05, offset=99999)
    return State(pad_periodic(u), pad_periodic(v), pad_periodic(w), ghosted_zeros(nx, ny, nz), _ghost_c(c))


def test_channel_cylinder_fft_line_solve_bitwise():
    o, g = run_both(configs.channel_cylinder(nx=64, ny=32, nz=16), 5)
    assert_bitwise(o, g)
    fo, fg = o.marker_forces(), g.marker_forces()
    for a, b in zip(fo, fg):
        assert np.array_equal(a, b)


def test_channel_cylinder_smac_spreading_weight_bitwise():
    """ibm_weight = smac.f90's dV = pi DC/NL dX^2 (smac.f90:1129, 1428) as
    the spreading and drag weight: FP64 bit-identical to the oracle, marker
    forces and drag included; the mixed step within 1e-5 of it."""
    from fortran8_amd.fields import cylinder_markers
    base = configs.channel_cylinder(nx=64, ny=32, nz=16)
    d = base.params.lx / base.params.nx
    nl = cylinder_markers([(1.0, 1.0)], 0.5, d, d, 16).nl
    dv = math.pi * 0.5 / nl * d * d
    case = configs.channel_cylinder(nx=64, ny=32, nz=16, ibm_weight=dv)
    o, g = run_both(case, 5)
    assert_bitwise(o, g)
    for a, b in zip(o.marker_forces(), g.marker_forces()):
        assert np.array_equal(a, b)
    do, dg = o.diagnostics(), g.diagnostics()
    # the same marker forces, summed in another order
    assert abs(dg["drag"] - do["drag"]) <= 1e-12 * abs(do["drag"]) and dg["drag"] != 0.0
    h = api.load_case(base)
    h.step(5)
    assert dg["drag"] != h.diagnostics()["drag"]   # the weight is applied
    mc = configs.channel_cylinder(nx=64, ny=32, nz=16, ibm_weight=dv, precision=1)
    m = api.load_case(mc)
    m.step(5)
    sm, sg = m.get_state(), g.get_state()
    # the velocity stacked (V and W alone are the initial noise, 1e-3 of U)
    um = np.concatenate([interior(getattr(sm, n)).ravel() for n in "UVW"])
    ug = np.concatenate([interior(getattr(sg, n)).ravel() for n in "UVW"])
    assert rel_l2(um, ug) < 1e-5
    assert rel_l2(interior_c(sm.C), interior_c(sg.C)) < 1e-5
    assert abs(m.diagnostics()["drag"] - dg["drag"]) < 1e-5 * abs(dg["drag"])


def test_channel_cylinder_sor_variant_at_solver_tolerance():
    case = configs.channel_cylinder(nx=32, ny=16, nz=8, poisson_variant=POISSON_SOR)
    case.params.sor_eps = 1e-10
    case.params.sor_itrmax = 20000
    o, g = run_both(case, 3)
    assert_close(o, g, 1e-8)
    assert g.diagnostics()["sor_itr"] > 0


def test_throughput_layout_small_bitwise():
    o, g = run_both(configs.throughput(nx=64, ny=64, nz=16, ncyl=2), 4)
    assert_bitwise(o, g)


def test_periodic_ibm_adjoint_config_bitwise():
    o, g = run_both(configs.adjoint_case(nx=32, ny=32, nz=16), 4)
    assert_bitwise(o, g)


def test_explicit_method_sor_at_solver_tolerance():
    case = configs.kolmogorov(n=16)
    case.params.method = METHOD_EXPLICIT
    case.params.poisson_variant = POISSON_SOR
    case.params.sor_eps = 1e-10
    case.params.sor_itrmax = 20000
    o, g = run_both(case, 3)
    assert_close(o, g, 1e-8)


def test_restart_through_the_boundary_is_exact():
    case = configs.kolmogorov(n=16)
    g1 = api.load_case(case)
    g1.step(6)
    g2 = api.load_case(case)
    g2.step(3)
    st = g2.get_state()
    g3 = api.FenepGPU(case.params)
    g3.set_state(st)
    g3.set_body_force(case.force)
    g3.step(3)
    a, b = g1.get_state(), g3.get_state()
    for name in FIELDS:
        assert np.array_equal(getattr(a, name), getattr(b, name))
    assert np.array_equal(a.C, b.C)


def test_diagnostics_agree_with_oracle():
    o, g = run_both(configs.channel_cylinder(nx=32, ny=16, nz=8), 3)
    do, dg = o.diagnostics(), g.diagnostics()
    for k in ("ke", "trc_min", "trc_max", "trc_mean", "drag", "lift"):
        assert dg[k] == pytest.approx(do[k], rel=1e-12, abs=1e-14), k
    assert dg["div_max"] < 1e-10 and dg["nan_flag"] == 0
    assert dg["newton_itr_mean"] == pytest.approx(do["newton_itr_mean"], rel=1e-12)


def test_full_size_throughput_case_two_steps_bitwise():
    case = configs.throughput()
    o, g = run_both(case, 2)
    assert_bitwise(o, g)


def test_full_size_throughput_properties_over_twenty_steps():
    case = configs.throughput()
    g = api.load_case(case)
    g.step(20)
    d = g.diagnostics()
    assert d["nan_flag"] == 0
    assert d["div_max"] < 1e-10
    assert 2.9 < d["trc_min"] and d["trc_max"] < 400.0
    assert d["cfl"] < 1.0


def test_nan_is_reported():
    case = configs.taylor_green(n=8)
    case.state.U[3, 3, 3] = np.nan
    g = api.load_case(case)
    g.step(1)
    assert g.diagnostics()["nan_flag"] == 1


def test_error_paths():
    p = default_params(nx=8, ny=8, nz=8)
    g = api.FenepGPU(p)
    with pytest.raises(api.FenepError, match="FP_E_STATE"):
        g.step(1)
    case = configs.channel_cylinder(nx=32, ny=16, nz=8)
    g = api.FenepGPU(case.params)
    g.set_state(case.state)
    with pytest.raises(api.FenepError, match="FP_E_STATE"):
        g.step(1)   # IBM method without markers
    bad = cylinder_markers([(1.0, 0.05)], 0.5, 0.0625, 0.0625, 8)
    with pytest.raises(api.FenepError, match="FP_E_ARG"):
        g.set_markers(bad)


def test_device_pointer_boundary_round_trip():
    import os
    import subprocess
    import sys
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_devptr_check.py")
    r = subprocess.run([sys.executable, script], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "round trip ok" in r.stdout


def test_time_stats_match_oracle():
    """all_time_dif statistics reduced on the device equal the oracle's."""
    for case in (configs.kolmogorov(n=32), configs.channel_cylinder(nx=64, ny=32, nz=16)):
        print("time stats", case.name, flush=True)
        o, g = po.load_case(case), api.load_case(case)
        o.step(3)
        g.step(3)
        so, sg = o.time_stats(), g.time_stats()
        np.testing.assert_allclose(sg, so, rtol=1e-12, atol=1e-300)
        g.close()
        o.close()


def test_energy_spectrum_matches_oracle():
    case = configs.kolmogorov(n=32)
    print("energy spectrum", case.name, flush=True)
    o, g = po.load_case(case), api.load_case(case)
    o.step(3)
    g.step(3)
    Eo, Eg = o.energy_spectrum(), g.energy_spectrum()
    np.testing.assert_allclose(Eg, Eo, rtol=1e-11, atol=1e-14 * Eo.sum())
    g.close()
    o.close()


def _close_with_nan(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    np.testing.assert_allclose(a[ok], b[ok], rtol=rtol, atol=1e-300)


def test_map_statistics_and_z_spectra_match_oracle_with_and_without_mean():
    """all_time_dif_map and fft_energy_dif_map on the device against the
    oracle, for the raw velocity ('same') and after the z-averaged time mean
    of three states is subtracted ('diff'), periodic and with walls."""
    for case, centres in ((configs.kolmogorov(n=32, nz=16), [(1.0, 2.0), (4.0, 4.5)]),
                          (configs.channel_cylinder(nx=64, ny=32, nz=16), [(2.0, 2.0)])):
        print("map statistics", case.name, flush=True)
        o, g = po.load_case(case), api.load_case(case)
        for _ in range(3):
            o.step(1)
            g.step(1)
            o.mean_accumulate()
            g.mean_accumulate()
        for mode in (0, 1):
            if mode:
                o.mean_apply(1)
                g.mean_apply(1)
            rows_o, rec_o = o.time_stats_map(centres, 0.5)
            rows_g, rec_g = g.time_stats_map(centres, 0.5)
            _close_with_nan(rows_g, rows_o, 1e-11)
            _close_with_nan(rec_g, rec_o, 1e-11)
            np.testing.assert_allclose(g.time_stats(), o.time_stats_np(), rtol=1e-11, atol=1e-300)
            em_o, ef_o = o.energy_spectrum_z(centres, 0.5)
            em_g, ef_g = g.energy_spectrum_z(centres, 0.5)
            np.testing.assert_allclose(em_g, em_o, rtol=1e-9, atol=1e-13 * em_o.sum())
            np.testing.assert_allclose(ef_g, ef_o, rtol=1e-9, atol=1e-13 * max(ef_o.sum(), 1e-300))
            if not case.params.walls_y:
                Eo, Eg = o.energy_spectrum(), g.energy_spectrum()
                if mode:
                    # the oracle's shell spectrum of the fluctuation
                    Eo = _shell_spectrum_of_fluct(o)
                np.testing.assert_allclose(Eg, Eo, rtol=1e-10, atol=1e-14 * Eo.sum())
        g.close()
        o.close()


def _shell_spectrum_of_fluct(o):
    st, (u, v, w) = o._fluct()
    nx, ny, nz = o.nx, o.ny, o.nz
    N = nx * ny * nz
    ureg = (u[1:-1, 1:-1, 1:-1] + u[1:-1, 1:-1, :-2]) / 2.0
    vreg = (v[1:-1, 1:-1, 1:-1] + v[1:-1, :-2, 1:-1]) / 2.0
    wreg = (w[1:-1, 1:-1, 1:-1] + w[:-2, 1:-1, 1:-1]) / 2.0
    e = sum(np.abs(np.fft.rfftn(a) / N) ** 2 for a in (ureg, vreg, wreg))
    kx = np.arange(nx // 2 + 1, dtype=float)[None, None, :]
    ky = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny).astype(float)[None, :, None]
    kz = np.where(np.arange(nz) <= nz // 2, np.arange(nz), np.arange(nz) - nz).astype(float)[:, None, None]
    idx = np.floor(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) + 0.5).astype(int)
    wgt = np.ones(nx // 2 + 1)
    wgt[0] = wgt[-1] = 0.5
    E = np.bincount(idx.ravel(), weights=(e * wgt[None, None, :]).ravel(), minlength=idx.max() + 1)
    return E[:nx] if len(E) >= nx else np.pad(E, (0, nx - len(E)))


def test_mean_accumulation_of_the_mixed_state():
    case = configs.kolmogorov(n=32, nz=16)
    case.params.precision = 1
    g = api.load_case(case)
    g.step(2)
    g.mean_accumulate()
    g.mean_apply(2)
    s = g.time_stats()
    assert s[2] == 0.0 and s[3] > 0.0
    with pytest.raises(api.FenepError, match="FP_E_STATE"):
        g.mean_apply(1)


@pytest.mark.parametrize("precision", [0, 1])
def test_setters_wait_for_steps_in_flight(precision):
    """fp_step(..., sync = 0) returns after enqueueing; every setter first
    waits for those steps.  A context that sets a state, a force and the
    markers right after an asynchronous run must continue exactly as a fresh
    context given the same inputs."""
    case = configs.channel_cylinder(nx=64, ny=32, nz=16, precision=precision)
    fresh = api.load_case(case)
    fresh.step(3)
    ref = fresh.get_state()
    g = api.load_case(case)
    g.step(40, sync=False)            # still running while the setters below are called
    g.set_state(case.state)
    g.set_body_force(case.force)
    g.set_markers(case.markers)
    g.step(3)
    got = g.get_state()
    for name in FIELDS:
        assert np.array_equal(interior(getattr(got, name)), interior(getattr(ref, name))), name
    assert np.array_equal(interior_c(got.C), interior_c(ref.C))
