"""The certified FP32 C+- selection of the mixed path (m_select in
fortran8_amd/csrc/fp_kernels_mixed.hip): the psi polynomial and the error-bound
constants read from the kernel source, checked on the CPU.

psi(s) = cos(acos(1 - s^2)/3 + 2 pi/3) is checked against numpy's arccos/cos
on [0, sqrt 2].  The error bound is checked on a float32 model of the kernel's
arithmetic (the same operation order, fmaf as one rounding, the reciprocal
square root with one-ulp perturbations) against the reference's Vieta formula
(fenep.f90:425-444) in float64 on the same inputs: the bound must exceed the
discrepancy with a margin, and every input where the reference returns its
sentinel (or NaN) must be flagged.  The GPU test
test_gpu_mixed.py::test_mixed_selection_is_the_reference_selection checks the
kernel itself against the FP64 selection."""
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fortran8_amd", "csrc",
                   "fp_kernels_mixed.hip")
f32, f64 = np.float32, np.float64
EPS32, EPS64 = 2.0 ** -24, 2.0 ** -53


def constants():
    txt = open(SRC).read()
    c = {k: float(v) for k, v in re.findall(r"#define (PSI\d|SEL_K\w+) ([-+0-9.eE]+)f", txt)}
    return [c["PSI%d" % i] for i in range(9)], c


def psi_exact(s):
    return np.cos(np.arccos(np.clip(1 - s * s, -1, 1)) / 3 + 2 * np.pi / 3)


def test_psi_polynomial_matches_arccos_cos():
    psi, _ = constants()
    s = np.linspace(0, np.sqrt(2), 100001)
    p = np.polynomial.polynomial.polyval(s, psi)
    # below half an FP32 ulp of |psi| >= 0.5 (2^-25), coefficients as printed
    assert np.max(np.abs(p - psi_exact(s))) < 2.0 ** -25
    # the two ends: a double smallest root (s = 0) and a double largest root
    assert psi[0] == -0.5
    assert abs(np.polynomial.polynomial.polyval(np.sqrt(2), psi) + 1) < 2.0 ** -25


def test_sel_constants_are_the_stated_multiples_of_eps():
    _, c = constants()
    assert c["SEL_K1E"] == np.float32(8 * EPS32)
    assert c["SEL_K2E"] == np.float32(8 * EPS32)
    assert c["SEL_KME"] == np.float32(8 * EPS32)
    assert c["SEL_KCE"] == np.float32(8 * EPS32)
    assert c["SEL_K3E"] == np.float32(64 * EPS64)


class Model:
    """float32 model of lmin32 (kernel order), one-ulp noise on the
    hardware reciprocal square root."""

    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)
        self.psi, c = constants()
        self.k1, self.k2, self.k3, self.km = c["SEL_K1E"], c["SEL_K2E"], c["SEL_K3E"], c["SEL_KME"]

    def noise(self, x):
        return (x * (1 + EPS32 * 2 * (self.rng.random(x.shape) - 0.5))).astype(f32)

    @staticmethod
    def fma(a, b, c):
        return (np.asarray(a, f64) * np.asarray(b, f64) + np.asarray(c, f64)).astype(f32)

    def lmin(self, a):
        fma = self.fma
        third = f32(1 / 3)
        m = ((a[0] + a[3]) + a[5]) * third
        b11, b22, b33 = a[0] - m, a[3] - m, a[5] - m
        o2 = fma(a[1], a[1], fma(a[2], a[2], a[4] * a[4]))
        hp = fma(b11, b11, fma(b22, b22, b33 * b33))
        r2 = fma(hp, f32(1 / 6), o2 * third)
        t1 = fma(b22, b33, -(a[4] * a[4]))
        t2 = fma(a[1], b33, -(a[4] * a[2]))
        t3 = fma(a[1], a[4], -(b22 * a[2]))
        det = fma(b11, t1, fma(-a[1], t2, a[2] * t3))
        with np.errstate(all="ignore"):
            ir = self.noise((1 / np.sqrt(r2)).astype(f32))
            r = r2 * ir
            ir3 = (ir * ir) * ir
            arg = (f32(0.5) * det) * ir3
            s2 = f32(1) - arg
            rho = fma(np.abs(m), ir, f32(2))
            d = fma(rho, fma(f32(self.k3), rho * rho, f32(self.k2)), f32(self.k1))
            v = self.noise((1 / np.sqrt(np.fmax(s2, d))).astype(f32))
            s = s2 * v
            # 2 psi with the coefficients doubled; lambda = m + r (2 psi)
            ps = np.full_like(s, f32(2 * self.psi[8]))
            for c in self.psi[7::-1]:
                ps = fma(ps, s, f32(2 * c))
            lam = fma(r, ps, m)
            amb = ~(np.abs(arg) <= fma(f32(-2), d, f32(1)))
            # the caller adds the cells' entry term (0 here)
            err = r * fma(d, v, f32(self.km) * rho)
        return lam.astype(f64), err.astype(f64), amb


def vieta_ref(a):
    """fenep.f90:425-444 in float64, as written"""
    a = a.astype(f64)
    with np.errstate(all="ignore"):
        a2 = -((a[0] + a[3]) + a[5])
        a1 = (((((a[0] * a[3]) + (a[3] * a[5])) + (a[5] * a[0])) - (a[4] * a[4])) - (a[2] * a[2])) - (a[1] * a[1])
        a0 = -((((((a[0] * a[3]) * a[5] + (a[1] * a[4]) * a[2]) + (a[1] * a[4]) * a[2]) - (a[0] * a[4]) * a[4]) -
                (a[1] * a[1]) * a[5]) - (a[2] * a[3]) * a[2])
        p = a1 - (a2 * a2) / 3.0
        q = (a0 - (a1 * a2) / 3.0) + (2.0 * ((a2 * a2) * a2)) / 27.0
        t = (q / 2.0) ** 2 + ((p / 3.0) * (p / 3.0)) * (p / 3.0)
        arg = ((3.0 * q) / (2.0 * p)) * np.sqrt(-(3.0 / p))
        x = (2.0 * np.sqrt(-(p / 3.0))) * np.cos(np.arccos(arg) / 3.0 + (2.0 * np.pi) / 3.0) - a2 / 3.0
    return np.where(t >= 0, -100.0, x)


def sym(rng, lam):
    q, _ = np.linalg.qr(rng.standard_normal((lam.shape[0], 3, 3)))
    m = q @ (lam[:, :, None] * np.transpose(q, (0, 2, 1)))
    return np.stack([m[:, 0, 0], m[:, 0, 1], m[:, 0, 2], m[:, 1, 1], m[:, 1, 2], m[:, 2, 2]]).astype(f32)


def families(rng, n):
    sc = 10 ** rng.uniform(-1, 2, (n, 1))
    base = sc * rng.uniform(0.05, 0.3, (n, 1))
    big = sc * rng.uniform(0.5, 1, (n, 1))
    gap = 10 ** rng.uniform(-8, -1, (n, 1))
    dev = 10 ** rng.uniform(-7, 0, (n, 1))
    uni = np.zeros((6, n), f32)
    uni[0] = (sc[:, 0] * rng.uniform(1, 50, n)).astype(f32)
    uni[3] = uni[5] = sc[:, 0].astype(f32)
    return {
        "definite": sym(rng, sc * rng.uniform(0.01, 1, (n, 3))),
        "indefinite": sym(rng, sc * rng.uniform(-1, 1, (n, 3))),
        "double smallest": sym(rng, np.concatenate([base, base * (1 + gap), big], 1)),
        "double largest": sym(rng, np.concatenate([base, big, big * (1 + gap)], 1)),
        "near isotropic": sym(rng, sc * (1 + dev * rng.uniform(-1, 1, (n, 3)))),
        "uniaxial": uni,
    }


def test_error_bound_covers_fp32_against_reference_formula():
    model = Model(3)
    rng = np.random.default_rng(4)
    for name, a in families(rng, 40000).items():
        lam, err, amb = model.lmin(a)
        ref = vieta_ref(a)
        sentinel = (ref == -100.0) | np.isnan(ref)
        assert not np.any(~amb & sentinel), name
        ok = ~amb & ~sentinel
        ratio = np.max(np.abs(lam[ok] - ref[ok]) / err[ok]) if ok.any() else 0.0
        assert ratio < 0.5, (name, ratio)


def test_bound_is_not_vacuous_on_well_separated_spectra():
    """Generic definite matrices are certified (not flagged) and their bound
    is a small multiple of eps32 times the spectral scale."""
    model = Model(5)
    rng = np.random.default_rng(6)
    a = sym(rng, rng.uniform(1, 10, (20000, 3)))
    lam, err, amb = model.lmin(a)
    assert amb.mean() < 0.01
    S = np.abs(a.astype(f64)).sum(0)
    assert np.median(err[~amb] / S[~amb]) < 1e-5
