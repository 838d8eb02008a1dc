"""Pins of the oracle's step composition by exact properties of the scheme
(CPU only).  Each test is built so that a plausible restatement mistake in the
part it pins fails it:

- CStar (fenep.f90:464-507): a smooth conformation carried by a uniform flow
  must converge to the exactly shifted field, along each axis and in both
  directions.  Taking the downwind face reconstruction (Cp and Cn exchanged)
  is unstable and fails.
- The predictor's trapezoid in the polymer stress and the Adams-Bashforth
  history (fenep.f90:803-806, 1469-1474): a uniform change of a history field
  moves the next velocity by exactly dt/2 (T0) or -dt/2 (A0, and B0 in the
  explicit method) times that change, and a uniform body force by dt.
- The Crank-Nicolson pressure update P += Phi - (dt beta / 2 Re) lap Phi
  (fenep.f90:1592-1595) and the explicit P += Phi (fenep.f90:937): a body
  force that is a discrete gradient, F = grad_h phi, from rest must be
  balanced by the pressure in one step, P = phi and u = 0, and stay there.
"""
import math

import numpy as np
import pytest

from fortran8_amd import configs
from fortran8_amd.fields import State, pad_periodic, interior, interior_c
from fortran8_amd.params import default_params, METHOD_FFT, METHOD_EXPLICIT, POISSON_FFT
from oracle import pyoracle as po

TWO_PI = 2 * math.pi


# ------------------------------------------------------------------ CStar
def conformation_profile(s):
    """SPD conformation field along one coordinate s (any axis): a fixed
    anisotropic part plus smooth periodic variations of every component."""
    c = np.zeros(s.shape + (6,))
    c[..., 0] = 2.0 + 0.6 * np.sin(s)
    c[..., 1] = 0.3 * np.cos(s)
    c[..., 2] = 0.2 * np.sin(2 * s)
    c[..., 3] = 1.2 + 0.4 * np.cos(s)
    c[..., 4] = -0.1 * np.sin(s)
    c[..., 5] = 0.8 + 0.2 * np.sin(s + 1.0)
    return c


def transport_error(axis, sign, n, t_end=1.0, cfl=0.1):
    """Carry conformation_profile along `axis` with a uniform velocity of
    `sign` for t_end; return the max error against the shifted profile.

    Wi = 1e12 makes the relaxation dt/Wi negligible and the polymer stress
    coefficient (1-beta)/(Re Wi) vanish, so the uniform velocity stays
    uniform and the velocity gradient is zero: C^{n+1} = C*(C^n), the CStar
    transport alone."""
    shape = [4, 4, 4]                # (nz, ny, nx)
    shape[2 - axis] = n
    nz, ny, nx = shape
    d = TWO_PI / n
    dt = cfl * d
    steps = int(round(t_end / dt))
    sc = (np.arange(1, n + 1) - 0.5) * d          # cell centres along the axis
    s = np.broadcast_to(sc.reshape([-1 if a == 2 - axis else 1 for a in range(3)]), (nz, ny, nx))
    C0 = conformation_profile(s)
    vel = [np.zeros((nz, ny, nx)) for _ in range(3)]
    vel[axis][:] = sign
    p = default_params(nx=nx, ny=ny, nz=nz, lx=nx * d, ly=ny * d, lz=nz * d, dt=dt, re=1.0, beta=0.5,
                       wi=1.0e12, lp=10.0, method=METHOD_FFT)
    st = State(*(pad_periodic(v) for v in vel), pad_periodic(np.zeros((nz, ny, nx))),
               np.pad(C0, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap"))
    o = po.Oracle(p)
    o.set_state(st)
    o.step(steps)
    out = o.get_state()
    for a, nm in enumerate("UVW"):
        assert np.max(np.abs(interior(getattr(out, nm)) - vel[a])) < 1e-12, "velocity stayed uniform"
    exact = conformation_profile(s - sign * steps * dt)
    return np.max(np.abs(interior_c(out.C) - exact))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_cstar_transport_converges_to_the_shifted_field(axis, sign):
    errs = [transport_error(axis, sign, n) for n in (32, 64, 128)]
    print("axis", axis, "sign", sign, "max errors", errs)
    # the upwind reconstruction converges at (at least) first order
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]
    assert errs[2] < 0.02


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_cstar_takes_the_upwind_face_only(axis, sign):
    """Domain of dependence of one step under a uniform flow of `sign`: the
    new C of cell i reads the face values of cells i and i - sign, whose
    reconstructions reach cells i - 2 sign .. i + sign.  A change at cell
    i + 2 sign (two cells downstream) must leave cell i bit-identical, and a
    change at i - 2 sign (upstream) must reach it."""
    n = 16
    shape = [4, 4, 4]
    shape[2 - axis] = n
    nz, ny, nx = shape
    d = TWO_PI / n
    sc = (np.arange(1, n + 1) - 0.5) * d
    s = np.broadcast_to(sc.reshape([-1 if a == 2 - axis else 1 for a in range(3)]), (nz, ny, nx))
    C0 = conformation_profile(s)
    vel = [np.zeros((nz, ny, nx)) for _ in range(3)]
    vel[axis][:] = sign
    p = default_params(nx=nx, ny=ny, nz=nz, lx=nx * d, ly=ny * d, lz=nz * d, dt=0.1 * d, re=1.0, beta=0.5,
                       wi=1.0e12, lp=10.0, method=METHOD_FFT)

    def new_c(C):
        o = po.Oracle(p)
        o.set_state(State(*(pad_periodic(v) for v in vel), pad_periodic(np.zeros((nz, ny, nx))),
                          np.pad(C, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap")))
        o.step(1)
        return np.moveaxis(interior_c(o.get_state().C), 2 - axis, 0)   # axis first

    i = 8
    base = new_c(C0)
    for offset, reaches in ((2 * int(sign), False), (-2 * int(sign), True)):
        C = np.moveaxis(C0.copy(), 2 - axis, 0)
        C[i + offset, ..., 0] += 0.05
        C[i + offset, ..., 3] += 0.03
        got = new_c(np.moveaxis(C, 0, 2 - axis))
        if reaches:
            assert np.max(np.abs(got[i] - base[i])) > 1e-6
        else:
            assert np.array_equal(got[i], base[i]), "cell i read a face value from downstream"


def test_cstar_without_flow_leaves_a_varying_conformation_unchanged():
    """u = 0: each cell's own face reconstructions Cp + Cn = 2C carry
    weight 1/6 per axis, so C* = C and the step is the pointwise Peterlin
    relaxation of each cell, whatever the neighbours hold."""
    n = 8
    rng = np.random.default_rng(7)
    s = rng.uniform(0, TWO_PI, (n, n, n))
    C = conformation_profile(s)
    z = np.zeros((n, n, n))
    dt, wi, lp = 0.01, 0.5, 10.0
    p = default_params(nx=n, ny=n, nz=n, dt=dt, re=1.0, beta=0.5, wi=wi, lp=lp)
    o = po.Oracle(p)
    o.set_state(State(pad_periodic(z), pad_periodic(z), pad_periodic(z), pad_periodic(z),
                      np.pad(C, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap")))
    o.step(1)
    got = interior_c(o.get_state().C)
    # (1 + dt f(x)/Wi) x = C + dt/Wi I per cell, f = (lp^2-3)/(lp^2 - tr x)
    x = C.copy()
    eye = np.array([1.0, 0, 0, 1.0, 0, 1.0])
    for _ in range(100):
        f = (lp * lp - 3) / (lp * lp - (x[..., 0] + x[..., 3] + x[..., 5]))
        x = (C + dt / wi * eye) / (1 + dt * f / wi)[..., None]
    assert np.max(np.abs(got - x)) < 1e-8


# ------------------------------------------------------ predictor weights
def developed_state(method):
    """A Kolmogorov FENE-P state with histories after three oracle steps."""
    case = configs.kolmogorov(n=16)
    case.params.method = method
    o = po.load_case(case)
    o.step(3)
    return case, o.get_state()


def one_step_from(case, st, body_force=None):
    o = po.Oracle(case.params)
    o.set_state(st)
    o.set_body_force(case.force if body_force is None else body_force)
    o.step(1)
    return o.get_state()


@pytest.mark.parametrize("method,hist_index,weight", [
    (METHOD_FFT, 3, 0.5),        # T0x: trapezoid dt (T + T0)/2
    (METHOD_FFT, 4, 0.5),        # T0y
    (METHOD_FFT, 0, -0.5),       # A0x: dt (3A - A0)/2
    (METHOD_EXPLICIT, 5, 0.5),   # T0z
    (METHOD_EXPLICIT, 1, -0.5),  # A0y: dt (3(A+B) - (A0+B0))/2
    (METHOD_EXPLICIT, 8, -0.5),  # B0z
])
def test_uniform_history_change_moves_the_velocity_by_its_weight(method, hist_index, weight):
    case, st = developed_state(method)
    base = one_step_from(case, st)
    pert = st.copy()
    delta = 0.7
    pert.hist[hist_index] += delta
    moved = one_step_from(case, pert)
    comp = hist_index % 3
    dt = case.params.dt
    for a, nm in enumerate("UVW"):
        du = interior(getattr(moved, nm)) - interior(getattr(base, nm))
        expect = weight * dt * delta if a == comp else 0.0
        assert np.max(np.abs(du - expect)) < 1e-12, (nm, du.mean(), expect)
    # the pressure and the conformation do not feel a uniform history shift
    assert np.max(np.abs(interior(moved.P) - interior(base.P))) < 1e-11
    assert np.array_equal(interior_c(moved.C), interior_c(base.C))


@pytest.mark.parametrize("hist_index,weight", [(3, 0.5), (5, 0.5), (1, -0.5)])
def test_ibm_predictor_history_weights_with_shifted_targets(hist_index, weight):
    """IBM method (ibm_preliminary, fenep.f90:1997-2000): the direct forcing
    is Galilean-invariant, so a uniform history change together with the
    same shift of every marker's target velocity leaves the marker forces
    unchanged (the delta kernel sums to one) and moves the new velocity by
    weight * dt * delta everywhere."""
    case = configs.adjoint_case(nx=16, ny=16, nz=8)
    o = po.load_case(case)
    o.step(3)
    st = o.get_state()
    dt, delta = case.params.dt, 0.7
    comp = hist_index % 3
    base = po.load_case(case)
    base.set_state(st)
    base.step(1)
    mk = case.markers
    shifted = [mk.uc, mk.vc, mk.wc][comp] + weight * dt * delta
    import dataclasses
    mk2 = dataclasses.replace(mk, **{"uvw"[comp] + "c": shifted})
    pert = st.copy()
    pert.hist[hist_index] += delta
    moved = po.Oracle(case.params)
    moved.set_state(pert)
    moved.set_body_force(case.force)
    moved.set_markers(mk2)
    moved.step(1)
    sb, sm = base.get_state(), moved.get_state()
    for a, nm in enumerate("UVW"):
        du = interior(getattr(sm, nm)) - interior(getattr(sb, nm))
        expect = weight * dt * delta if a == comp else 0.0
        assert np.max(np.abs(du - expect)) < 1e-12, (nm, du.mean(), expect)
    for fa, fb in zip(base.marker_forces(), moved.marker_forces()):
        assert np.max(np.abs(fa - fb)) < 1e-9


@pytest.mark.parametrize("method", [METHOD_FFT, METHOD_EXPLICIT])
def test_uniform_body_force_moves_the_velocity_by_dt(method):
    case, st = developed_state(method)
    base = one_step_from(case, st)
    fx, fy, fz = case.force
    moved = one_step_from(case, st, body_force=(fx, fy + 0.3, fz))
    dv = interior(moved.V) - interior(base.V)
    assert np.max(np.abs(dv - 0.3 * case.params.dt)) < 1e-12


# ------------------------------------------------------- pressure update
def gradient_force(phi, d):
    """Staggered discrete gradient of a cell-centred periodic phi: the
    force sits where U, V, W sit (face i between cells i and i+1)."""
    return tuple((np.roll(phi, -1, axis=2 - a) - phi) / d[a] for a in range(3))


@pytest.mark.parametrize("method,re,dt", [(METHOD_FFT, 1.0, 0.1), (METHOD_FFT, 50.0, 0.02),
                                          (METHOD_EXPLICIT, 1.0, 0.01)])
def test_gradient_body_force_is_balanced_by_the_pressure_in_one_step(method, re, dt):
    nx, ny, nz = 16, 8, 8
    lx, ly, lz = TWO_PI, 3.0, 2.0
    d = (lx / nx, ly / ny, lz / nz)
    xc = (np.arange(nx) + 0.5) * d[0]
    yc = (np.arange(ny) + 0.5) * d[1]
    zc = (np.arange(nz) + 0.5) * d[2]
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    phi = np.sin(X) * np.cos(TWO_PI * Y / ly) + 0.3 * np.cos(2 * TWO_PI * Z / lz) + 0.2 * np.sin(2 * X + TWO_PI * Y / ly)
    phi -= phi.mean()
    z = np.zeros((nz, ny, nx))
    p = default_params(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, dt=dt, re=re, beta=1.0, method=method,
                       poisson_variant=POISSON_FFT)
    o = po.Oracle(p)
    o.set_state(State(pad_periodic(z), pad_periodic(z), pad_periodic(z), pad_periodic(z), np.zeros((nz + 2, ny + 2, nx + 2, 6))))
    o.set_body_force(gradient_force(phi, d))
    scale = np.max(np.abs(phi))
    for step in range(3):
        o.step(1)
        s = o.get_state()
        P = interior(s.P)
        assert np.max(np.abs((P - P.mean()) - phi)) < 1e-12 * scale, ("step", step + 1)
        for nm in "UVW":
            assert np.max(np.abs(interior(getattr(s, nm)))) < 1e-13, (nm, step + 1)
