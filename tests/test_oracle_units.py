"""Pins of the CPU oracle's per-cell helpers against known answers, library
routines and closed forms (CPU only)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def sym(c6):
    a = np.asarray(c6)
    return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])


def rand_spd(rng, scale=1.0, spread=0.5):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    lam = scale * (1.0 + spread * rng.random(3))
    m = (q * lam) @ q.T
    return np.array([m[0, 0], m[0, 1], m[0, 2], m[1, 1], m[1, 2], m[2, 2]]), lam.min()


# ------------------------------------------------------------ Vieta (A5)
def test_vieta_matches_eigvalsh_on_random_spd():
    rng = np.random.default_rng(0)
    for _ in range(500):
        c6, lmin = rand_spd(rng, scale=rng.uniform(0.1, 10.0))
        assert po.vieta(c6) == pytest.approx(lmin, rel=1e-9, abs=1e-12)


def test_vieta_indefinite_matrix_gives_its_smallest_eigenvalue():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.standard_normal(6)
        lam = np.linalg.eigvalsh(sym(a))
        if min(np.diff(lam)) < 1e-3:
            continue
        assert po.vieta(a) == pytest.approx(lam[0], rel=1e-8, abs=1e-10)


def test_vieta_diagonal_closed_form():
    assert po.vieta([3.0, 0, 0, 2.0, 0, 5.0]) == pytest.approx(2.0, abs=1e-13)
    assert po.vieta([7.0, 0, 0, -1.0, 0, 0.5]) == pytest.approx(-1.0, abs=1e-13)


def test_vieta_sentinel_on_repeated_eigenvalues():
    # exactly isotropic C = I: p = q = 0, t = 0 -> -100 (fenep.f90:438-439)
    assert po.vieta([1.0, 0, 0, 1.0, 0, 1.0]) == -100.0
    # the dsyev sample matrix has the double eigenvalue 1 (sample5.f90): the
    # reference's Viete test rejects it with the sentinel as well
    ka = KA["dsyev_3"]
    assert np.allclose(np.linalg.eigvalsh(sym(ka["C6"])), ka["eig"])
    assert po.vieta(ka["C6"]) == -100.0


# ---------------------------------------------------- selection (A5, C+-)
def brute_select(cm, c, cp):
    cand_p = [c * 3 / 2 - cp / 2, cm / 2 + c / 2, cm / 4 + c - cp / 4]
    cand_n = [c / 2 + cp / 2, -cm / 2 + c * 3 / 2, -cm / 4 + c + cp / 4]
    ml = []
    for l in range(3):
        ml.append(min(np.linalg.eigvalsh(sym(cand_p[l]))[0], np.linalg.eigvalsh(sym(cand_n[l]))[0]))
    ml.append(0.0)
    return int(np.argmax(ml)) + 1, ml, cand_p, cand_n


def test_selection_matches_eigvalsh_brute_force():
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(400):
        base, _ = rand_spd(rng)
        cm = base + 0.2 * rng.standard_normal(6)
        c = base + 0.2 * rng.standard_normal(6)
        cp = base + 0.2 * rng.standard_normal(6)
        idx_b, ml, cpl, cnl = brute_select(cm, c, cp)
        srt = sorted(ml, reverse=True)
        if srt[0] - srt[1] < 1e-6:
            continue  # near ties are decided by rounding; not comparable
        idx, po_, no_ = po.select(cm, c, cp)
        assert idx == idx_b
        if idx == 4:
            assert np.array_equal(po_, c) and np.array_equal(no_, c)
        else:
            assert np.allclose(po_, cpl[idx - 1], atol=1e-15)
            assert np.allclose(no_, cnl[idx - 1], atol=1e-15)
        checked += 1
    assert checked > 300


def test_selection_zero_slope_when_no_candidate_is_spd():
    cm = np.array([1.0, 0, 0, 1.0, 0, 1.0])
    c = np.array([0.05, 0, 0, 1.0, 0, 1.0])
    cp = np.array([-2.0, 0, 0, 1.0, 0, 1.0])
    idx, po_, no_ = po.select(cm, c, cp)
    assert idx == 4 and np.array_equal(po_, c) and np.array_equal(no_, c)


def test_selection_uniform_fields():
    # uniform anisotropic SPD: all candidates equal C, first maximum wins
    c = np.array([2.0, 0.1, 0.0, 1.5, 0.2, 1.0])
    idx, po_, no_ = po.select(c, c, c)
    assert idx == 1 and np.allclose(po_, c) and np.allclose(no_, c)
    # uniform isotropic C = I: every candidate hits the sentinel -> zero slope
    e = np.array([1.0, 0, 0, 1.0, 0, 1.0])
    assert po.select(e, e, e)[0] == 4


# ------------------------------------------------------------- gauss (A7)
@pytest.mark.parametrize("name", ["gauss_3", "gauss_4", "gauss_6"])
def test_gauss_known_answers(name):
    ka = KA[name]
    A = np.asarray(ka["A"])
    n = A.shape[0]
    M = np.eye(6)
    M[:n, :n] = A
    b = np.zeros(6)
    b[:n] = ka["b"]
    x = po.gauss6(M, b)
    assert np.allclose(x[:n], ka["x"], atol=1e-12)


def test_gauss_random_vs_numpy_and_zero_pivot():
    rng = np.random.default_rng(3)
    for _ in range(100):
        A = rng.standard_normal((6, 6))
        b = rng.standard_normal(6)
        assert np.allclose(po.gauss6(A, b), np.linalg.solve(A, b), rtol=1e-9, atol=1e-9)
    A = rng.standard_normal((6, 6))
    A[:, 2] = 0.0   # zero pivot column -> x = 0 (fenep.f90:621-625)
    assert np.array_equal(po.gauss6(A, rng.standard_normal(6)), np.zeros(6))


# ------------------------------------------------- Lyapunov Newton (A7)
def lyap_residual(A0T, b, x, dt, wi, lp):
    """(1 + dt f/Wi) X - dt (L X + X L^T) - C* - dt/Wi I in matrix form, with
    A = transpose(A0) = -dt L  (independent of the oracle's per-component
    expansion)."""
    X = sym(x)
    f = (lp * lp - 3) / (lp * lp - np.trace(X))
    A = np.asarray(A0T).reshape(3, 3, order="F")
    B = sym(b)  # b = C* + dt/Wi I
    R = (1 + dt * f / wi) * X + A @ X + X @ A.T - B
    return np.array([R[0, 0], R[0, 1], R[0, 2], R[1, 1], R[1, 2], R[2, 2]])


def test_newton_solves_the_implicit_conformation_equation():
    rng = np.random.default_rng(4)
    dt, wi, lp = 0.002, 2.0, 10.0
    for _ in range(50):
        A = -dt * rng.standard_normal((3, 3))
        cstar, _ = rand_spd(rng, scale=1.0, spread=0.3)
        b = cstar.copy()
        b[[0, 3, 5]] += dt / wi
        x, it = po.newton_cell(A.ravel(order="F"), b, cstar, dt, wi, lp)
        assert it <= 5
        assert np.max(np.abs(lyap_residual(A.ravel(order="F"), b, x, dt, wi, lp))) < 1e-12


def test_newton_zero_gradient_scalar_fixed_point():
    # L = 0: (1 + dt f(X)/Wi) X = C* + dt/Wi I with X = s I for C* = c I
    dt, wi, lp = 0.01, 1.0, 5.0
    c = 1.3
    b = np.array([c + dt / wi, 0, 0, c + dt / wi, 0, c + dt / wi])
    x, it = po.newton_cell(np.zeros(9), b, np.array([1.0, 0, 0, 1.0, 0, 1.0]), dt, wi, lp)
    s = x[0]
    f = (lp * lp - 3) / (lp * lp - 3 * s)
    assert (1 + dt * f / wi) * s == pytest.approx(c + dt / wi, rel=1e-12)
    assert abs(x[1]) < 1e-15 and x[3] == pytest.approx(s) and x[5] == pytest.approx(s)
    assert it <= 3


def test_newton_oldroyd_b_limit_is_a_linear_lyapunov_solve():
    # Lp -> infinity: f -> 1, the equation is linear: Kronecker-sum solve
    rng = np.random.default_rng(5)
    dt, wi, lp = 0.005, 1.0, 1.0e7
    A = -dt * rng.standard_normal((3, 3))
    cstar, _ = rand_spd(rng)
    b = cstar.copy()
    b[[0, 3, 5]] += dt / wi
    x, _ = po.newton_cell(A.ravel(order="F"), b, cstar, dt, wi, lp)
    a = (1 + dt / wi) / 2 * np.eye(3) + A
    K = np.kron(np.eye(3), a) + np.kron(a, np.eye(3))
    X = np.linalg.solve(K, sym(b).ravel(order="F")).reshape(3, 3, order="F")
    assert np.allclose(sym(x), X, rtol=1e-9, atol=1e-12)


# --------------------------------------------------------------- tdma (A20)
def test_tdma_known_answer():
    ka = KA["tdma_4"]
    x = po.tdma(ka["a"], ka["b"], ka["c"], ka["d"])
    assert np.allclose(x.real, ka["x"], atol=1e-14) and np.allclose(x.imag, 0)


def test_tdma_random_complex_rhs_vs_numpy():
    rng = np.random.default_rng(6)
    n = 37
    a = 4 + rng.random(n)
    b = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    M = np.diag(a) + np.diag(b[:-1], 1) + np.diag(c[1:], -1)
    assert np.allclose(po.tdma(a, b, c, d), np.linalg.solve(M, d), rtol=1e-12, atol=1e-12)


def test_tdma_singular_neumann_mode_pins_last_unknown():
    # Phi rows of fenep_2d.f90:1318-1329 for the zero mode: singular, the
    # 1e-16 guard sets the top ghost to 0 and the rest follows
    ny, dy = 8, 0.5
    n = ny + 2
    a = np.full(n, -2 / dy ** 2); b = np.full(n, 1 / dy ** 2); c = np.full(n, 1 / dy ** 2)
    a[0], b[0] = 1.0, -1.0
    a[-1], c[-1], b[-1] = 1.0, -1.0, 0.0
    rng = np.random.default_rng(7)
    d = np.zeros(n, complex)
    d[1:-1] = rng.standard_normal(ny)
    d[1:-1] -= d[1:-1].mean()
    x = po.tdma(a, b, c, d)
    assert x[-1] == 0
    lap = (x[:-2] - 2 * x[1:-1] + x[2:]) / dy ** 2
    assert np.allclose(lap, d[1:-1], atol=1e-11)
    assert x[0] == pytest.approx(x[1]) and x[-1] == pytest.approx(x[-2])


# ---------------------------------------------------------- Roma (A16)
def test_roma_kernel_moments_and_values():
    assert po.roma(0.0) == pytest.approx(2 / 3, abs=1e-15)
    assert po.roma(1.0) == pytest.approx(1 / 6, abs=1e-15)
    assert po.roma(1.5) == 0.0 and po.roma(-2.0) == 0.0
    assert po.roma(0.5 - 1e-12) == pytest.approx(po.roma(0.5), abs=1e-9)
    rng = np.random.default_rng(8)
    for x in rng.uniform(-3, 3, 200):
        pts = np.arange(math.floor(x) - 2, math.floor(x) + 3)
        w = np.array([po.roma(x - p) for p in pts])
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        assert (w * (x - pts)).sum() == pytest.approx(0.0, abs=1e-14)
        assert (w * w).sum() == pytest.approx(0.5, abs=1e-14)


# ---------------------------------------------------------------- FFT (A10)
@pytest.mark.parametrize("n", [4, 16, 128, 256])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(po.fft(x), np.fft.fft(x), rtol=1e-12, atol=1e-12 * n)
    assert np.allclose(po.fft(x, inverse=True), np.fft.ifft(x) * n, rtol=1e-12, atol=1e-12 * n)


def test_fft_naive_dft_small():
    n = 8
    x = np.arange(n) + 1j * np.arange(n)[::-1]
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.allclose(po.fft(x), dft, atol=1e-12)
