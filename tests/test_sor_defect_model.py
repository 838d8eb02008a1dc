"""The mixed-precision SOR variant on the CPU: a numpy float32 model of its
kernels (m_sor_start, m_sor_sweep, m_sor_accum and m_residual in
fp_kernels_mixed.hip) against FP64 red-black SOR of the same system, with the
reference's stopping test sqrt(sum E^2 / sum Q^2) < eps (poisson,
fenep.f90:848-905).

A smooth right-hand side from a zero start needs a correction e that grows
smooth and large over many sweeps.  Its FP32 stencil sum L_h e then cancels
down to the size of E, and the sweeps stall above the tolerance.  Moving e
into the FP64 potential every 10 sweeps and taking the defect again keeps the
FP32 part small: the sweeps then stop at the FP64 sweep count, with the FP64
residual below the tolerance and the potential within 1e-8 of FP64's."""
import numpy as np

NX, NY, NZ = 64, 32, 32
LX, LY, LZ = 8.0, 4.0, 4.0
DX, DY, DZ = LX / NX, LY / NY, LZ / NZ
BX, BY, BZ = 1 / DX ** 2, 1 / DY ** 2, 1 / DZ ** 2
B0 = ((((BX + BX) + BY) + BY) + BZ) + BZ
OM = 2 / (1 + np.sin(np.pi * DX))
EPS = 1e-5
K, J, I = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
COLOURS = [((I + J + K + c) & 1) == 0 for c in (0, 1)]


def ghost(a):
    """periodic x and z, Neumann walls in y (the boundary rule of the potential)"""
    g = np.pad(a, 1, mode="wrap")
    g[:, 0, :] = g[:, 1, :]
    g[:, -1, :] = g[:, -2, :]
    return g


def lap(a, dt=np.float64):
    """the sweep's stencil sum, in the kernel's order of operations"""
    g = ghost(a).astype(dt)
    c = g[1:-1, 1:-1, 1:-1]
    return (((((dt(BZ) * g[:-2, 1:-1, 1:-1] + dt(BY) * g[1:-1, :-2, 1:-1]) + dt(BX) * g[1:-1, 1:-1, :-2]) - dt(B0) * c)
             + dt(BX) * g[1:-1, 1:-1, 2:]) + dt(BY) * g[1:-1, 2:, 1:-1]) + dt(BZ) * g[2:, 1:-1, 1:-1]


def sweep(x, q, dt):
    """one red-black sweep; returns sum E^2 (E taken before each update)"""
    er = 0.0
    for msk in COLOURS:
        E = (lap(x, dt) - q.astype(dt)).astype(dt)
        x[msk] = (x[msk] + dt(OM / B0) * E[msk]).astype(dt)
        er += float(np.sum(E[msk].astype(np.float64) ** 2))
    return er


def smooth_rhs():
    x, y, z = (I + 0.5) * DX, (J + 0.5) * DY, (K + 0.5) * DZ
    b = np.cos(np.pi * y / LY) * (1 + 0.3 * np.cos(2 * np.pi * x / LX)) \
        + 0.2 * np.cos(2 * np.pi * z / LZ) * np.cos(2 * np.pi * y / LY)
    b = b + 1e-3 * np.random.default_rng(2).standard_normal(b.shape)
    return b - b.mean()


def fp32_defect_sweeps(b, anchor, cap):
    er0 = np.sum(b ** 2)
    phi = np.zeros_like(b)
    r = b - lap(phi)
    e = np.zeros(b.shape, np.float32)
    for it in range(1, cap + 1):
        done = np.sqrt(sweep(e, r.astype(np.float32), np.float32) / er0) < EPS
        if done:
            break
        if anchor and it % anchor == 0:
            phi = phi + e
            e[:] = 0
            r = b - lap(phi)
    phi = phi + e
    return it, done, phi


def test_fp32_sweeps_on_the_fp64_defect_stop_where_fp64_sor_stops():
    b = smooth_rhs()
    er0 = np.sum(b ** 2)
    x = np.zeros_like(b)
    for n64 in range(1, 10000):
        if np.sqrt(sweep(x, b, np.float64) / er0) < EPS:
            break
    it, done, phi = fp32_defect_sweeps(b, anchor=10, cap=2 * n64)
    assert done and it == n64, (it, n64)
    assert np.sqrt(np.sum((b - lap(phi)) ** 2) / er0) < EPS
    assert np.abs(phi - x).max() <= 1e-8 * np.abs(x).max()
    # without moving e into phi the FP32 sweeps stall above the tolerance
    it0, done0, _ = fp32_defect_sweeps(b, anchor=0, cap=int(1.5 * n64))
    assert not done0, it0
