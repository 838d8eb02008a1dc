"""The mixed-precision Lyapunov Newton near the FENE-P pole (CPU only).

A float32 model of the kernel's iteration (m_cstar_newton: analytic
Jacobian J = J0 + gp x e^T, LU without pivoting, chord steps, then full
Newton from the third iteration) on cells with 0.99 < tr C / L^2 < 0.999:

- every LU pivot stays above 1 + f dt/Wi less the size of the velocity
  gradient terms, and the float32 solve has a backward error at float32
  round-off (no growth from the missing pivoting);
- the iteration converges to the root of the FP64 oracle's Newton
  (fenep.f90:554-578, central-difference Jacobian, partial pivoting) to
  float32 accuracy, within the iteration cap.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def lyap_matrix(a, dg):
    """M(x) of the residual l = M x - b (the kernel's l[0..5]) as a 6x6
    matrix in float32, for the transposed velocity-gradient terms a (3x3)
    and dg = (1 + f dt/Wi)/2."""
    a11, a12, a13 = a[0]
    a21, a22, a23 = a[1]
    a31, a32, a33 = a[2]
    A11, A22, A33 = a11 + dg, a22 + dg, a33 + dg
    two = f32(2)
    z = f32(0)
    return np.array([[two * A11, two * a12, two * a13, z, z, z],
                     [a21, A22 + A11, a23, a12, a13, z],
                     [a31, a32, A33 + A11, z, a12, a13],
                     [z, two * a21, z, two * A22, two * a23, z],
                     [z, a31, a21, a32, A33 + A22, a23],
                     [z, z, two * a31, z, two * a32, two * A33]], dtype=f32)


def lu_nopivot(J):
    J = J.astype(f32).copy()
    piv = np.zeros(6, dtype=f32)
    for k in range(6):
        piv[k] = J[k, k]
        inv = f32(1) / J[k, k]
        for r in range(k + 1, 6):
            fct = J[r, k] * inv
            J[r, k + 1:] -= fct * J[k, k + 1:]
            J[r, k] = fct
        J[k, k] = inv
    return J, piv


def lu_solve(LU, l):
    l = l.astype(f32).copy()
    for k in range(6):
        for r in range(k + 1, 6):
            l[r] -= LU[r, k] * l[k]
    y = np.zeros(6, dtype=f32)
    for r in range(5, -1, -1):
        acc = l[r]
        for c in range(r + 1, 6):
            acc -= LU[r, c] * y[c]
        y[r] = acc * LU[r, r]
    return y


def near_pole_cell(rng, L2, frac, dtw, grad, ex=None):
    """An SPD conformation with tr C = frac L^2, stretched along a random
    direction n (x component ex if given: small values make the elimination
    multipliers x_r / x_0 large), and the transposed velocity-gradient terms
    a_ij = -dt du_i/dx_j (|.| <= grad)."""
    s = frac * L2
    n = rng.standard_normal(3)
    if ex is not None:
        n[0] = ex
    n /= np.linalg.norm(n)
    Q, _ = np.linalg.qr(np.column_stack([n, rng.standard_normal((3, 2))]))
    l2, l3 = rng.uniform(0.3, 1.0, 2)
    M = (s - l2 - l3) * np.outer(Q[:, 0], Q[:, 0]) + l2 * np.outer(Q[:, 1], Q[:, 1]) + l3 * np.outer(Q[:, 2], Q[:, 2])
    x = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])
    a = rng.uniform(-grad, grad, (3, 3))
    return x, a


CASES = [(L2, frac, dtw) for L2 in (9.0, 100.0, 400.0) for frac in (0.991, 0.995, 0.999)
         for dtw in (2.5e-4, 1e-3, 1e-2)]


@pytest.mark.parametrize("L2,frac,dtw", CASES)
def test_lu_without_pivoting_is_stable_near_the_pole(L2, frac, dtw):
    """For velocity gradients up to dt |du_i/dx_j| = 0.05 (u dt/dX is about
    0.03 in the throughput case, and CStar needs 6 CFL <= 1) the pivots stay
    above a quarter of 1 + f dt/Wi, the LU factors grow no larger than the
    matrix, and the float32 solve has a backward error at float32 round-off.
    (Measured beyond that range: at dt |du/dx| = 0.15 the growth reaches 20
    and the backward error 4e-7.)"""
    rng = np.random.default_rng(int(L2 * 1000 + frac * 1000 + dtw * 1e5))
    e = np.array([1, 0, 0, 1, 0, 1], dtype=f32)
    for ex in (None, 1e-3, 1e-2, 0.1):
        for _ in range(30):
            x, a = near_pole_cell(rng, L2, frac, dtw, 0.05, ex)
            s = x[0] + x[3] + x[5]
            f = (L2 - 3.0) / (L2 - s)
            gp = dtw * (L2 - 3.0) / (L2 - s) ** 2
            J = lyap_matrix(a.astype(f32), f32(0.5 * (1.0 + f * dtw))) + f32(gp) * np.outer(x.astype(f32), e)
            LU, piv = lu_nopivot(J)
            assert piv.min() >= 0.25 * (1.0 + f * dtw), (piv, f * dtw)
            assert np.abs(LU).max() <= 1.5 * np.abs(J).max()
            r = rng.standard_normal(6)
            y = lu_solve(LU, r).astype(np.float64)
            Jd = J.astype(np.float64)
            berr = np.linalg.norm(Jd @ y - r) / (np.linalg.norm(Jd, np.inf) * np.linalg.norm(y) + np.linalg.norm(r))
            assert berr < 2e-7, berr


def newton32(a, b, x0, dtw, L2, eps, itrmax):
    """Float32 model of the kernel's iteration: chord steps through the
    Jacobian of the first iterate, full Newton from the third iteration."""
    x = x0.astype(f32).copy()
    e = np.array([1, 0, 0, 1, 0, 1], dtype=f32)
    LU = None
    for it in range(1, itrmax + 1):
        s = float(x[0] + x[3] + x[5])
        f = (L2 - 3.0) / (L2 - s)          # FP64 within 1 % of L^2, as the kernel
        dg = f32(0.5 * (1.0 + f * dtw))
        M = lyap_matrix(a, dg)
        l = M @ x - b
        if it == 1 or it >= 3:
            gp = f32(dtw * (L2 - 3.0) / (L2 - s) ** 2)
            LU, _ = lu_nopivot(M + gp * np.outer(x, e))
        rr = lu_solve(LU, l)
        x = x - rr
        if float(rr @ rr) < eps * eps * float(x @ x):
            return x, it
    return x, -itrmax


@pytest.mark.parametrize("L2,frac,dtw", [c for c in CASES if c[2] >= 1e-3])
def test_chord_then_newton_converges_to_the_oracle_root_near_the_pole(L2, frac, dtw):
    rng = np.random.default_rng(int(L2 * 7 + frac * 1e4 + dtw * 1e6))
    dt = 0.002
    wi = dt / dtw
    lp = np.sqrt(L2)
    for _ in range(20):
        x0, a = near_pole_cell(rng, L2, frac, dtw, 0.03)
        # right-hand side C* + dt/Wi I with C* a small perturbation of C^n
        cstar = x0 * (1.0 + rng.uniform(-2e-3, 2e-3, 6))
        b = cstar + dtw * np.array([1, 0, 0, 1, 0, 1.0])
        # the oracle takes A(r, c) = a_rc column-major: A[r + 3 c]
        A = np.zeros(9)
        for r in range(3):
            for c in range(3):
                A[r + 3 * c] = a[r, c]
        xr, itr = po.newton_cell(A.reshape(3, 3, order="F"), b, x0, dt, wi, lp, eps=1e-12, itrmax=100)
        x32, it = newton32(a.astype(f32), b.astype(f32), x0, dtw, L2, 1e-6, 16)
        assert it > 0, "float32 iteration hit the cap"
        assert np.max(np.abs(x32 - xr)) <= 2e-5 * np.max(np.abs(xr)), (x32, xr)
