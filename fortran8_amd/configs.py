"""The five workload configurations, buildable at any grid size.

Each builder returns a ``Case`` with the parameters, the seeded initial state
(reference layout), an optional body-force field and optional IBM markers.
Reference-to-config parameter mapping: ``lp`` is L (L^2 = 100 -> lp = 10,
fenep.f90:513); ``re`` is U*1/nu in domain length units, so Re_D = 100 with
D = 0.5 gives re = 200.
"""
from dataclasses import dataclass
import math

import numpy as np

from .params import FpParams, default_params, METHOD_FFT, METHOD_IBM, POISSON_FFT, POISSON_SOR
from .fields import (State, Markers, seeded_uniform, seeded_normal, pad_periodic,
                     ghosted_zeros, cylinder_markers, seeded_cylinder_centres)


@dataclass
class Case:
    name: str
    params: FpParams
    state: State
    force: tuple = None          # (fx, fy, fz) interior arrays or None
    markers: Markers = None
    nsteps: int = 0              # the configuration's step count

    @property
    def cells(self):
        return self.params.nx * self.params.ny * self.params.nz


# ------------------------------------------------------------ projections
def _project_periodic(u, v, w, dx, dy, dz):
    """Make staggered (u at i, v at j, w at k faces) fields discretely
    divergence-free on a periodic grid with the exact inverse of the 7-point
    Laplacian (used for initial conditions only)."""
    nz, ny, nx = u.shape
    div = ((u - np.roll(u, 1, 2)) / dx + (v - np.roll(v, 1, 1)) / dy + (w - np.roll(w, 1, 0)) / dz)
    kx = 2 * (1 - np.cos(2 * np.pi * np.arange(nx) / nx)) / dx ** 2
    ky = 2 * (1 - np.cos(2 * np.pi * np.arange(ny) / ny)) / dy ** 2
    kz = 2 * (1 - np.cos(2 * np.pi * np.arange(nz) / nz)) / dz ** 2
    lap = -(kz[:, None, None] + ky[None, :, None] + kx[None, None, :])
    lap[0, 0, 0] = 1.0
    ph = np.fft.fftn(div) / lap
    ph[0, 0, 0] = 0
    phi = np.real(np.fft.ifftn(ph))
    u = u - (np.roll(phi, -1, 2) - phi) / dx
    v = v - (np.roll(phi, -1, 1) - phi) / dy
    w = w - (np.roll(phi, -1, 0) - phi) / dz
    return u, v, w


def _project_walls(u, v, w, dx, dy, dz):
    """Divergence-free projection with y-walls: v = 0 on j = 0 and j = NY
    faces, Neumann phi; cosine transform in y."""
    nz, ny, nx = u.shape
    v = v.copy()
    v[:, ny - 1, :] = 0.0
    vm = np.concatenate([np.zeros((nz, 1, nx)), v[:, :-1, :]], axis=1)
    div = (u - np.roll(u, 1, 2)) / dx + (v - vm) / dy + (w - np.roll(w, 1, 0)) / dz
    from scipy.fft import dctn, idctn
    kx = 2 * (1 - np.cos(2 * np.pi * np.arange(nx) / nx)) / dx ** 2
    ky = 2 * (1 - np.cos(np.pi * np.arange(ny) / ny)) / dy ** 2
    kz = 2 * (1 - np.cos(2 * np.pi * np.arange(nz) / nz)) / dz ** 2
    lap = -(kz[:, None, None] + ky[None, :, None] + kx[None, None, :])
    lap[0, 0, 0] = 1.0
    d = np.fft.fft(np.fft.fft(div, axis=2), axis=0)
    d = dctn(d.real, type=2, axes=[1], norm="ortho") + 1j * dctn(d.imag, type=2, axes=[1], norm="ortho")
    d = d / lap
    d[0, 0, 0] = 0
    d = idctn(d.real, type=2, axes=[1], norm="ortho") + 1j * idctn(d.imag, type=2, axes=[1], norm="ortho")
    phi = np.real(np.fft.ifft(np.fft.ifft(d, axis=2), axis=0))
    phip = np.concatenate([phi[:, 1:, :], phi[:, -1:, :]], axis=1)
    u = u - (np.roll(phi, -1, 2) - phi) / dx
    v = v - (phip - phi) / dy
    w = w - (np.roll(phi, -1, 0) - phi) / dz
    v[:, ny - 1, :] = 0.0
    return u, v, w


def _spd_perturbed_c(seed, nx, ny, nz, amp=1.0e-3, offset=0):
    """C = I + sym(U(-amp, amp)), eigenvalues clipped to >= 1e-6 (SPD)."""
    n = nx * ny * nz
    r = seeded_uniform(seed, 6 * n, offset=offset).reshape(nz, ny, nx, 6)
    c = 2.0 * amp * (r - 0.5)
    c[..., 0] += 1.0
    c[..., 3] += 1.0
    c[..., 5] += 1.0
    # Gershgorin: every eigenvalue >= min_r (c_rr - sum_{s!=r} |c_rs|)
    gersh = np.minimum(np.minimum(c[..., 0] - abs(c[..., 1]) - abs(c[..., 2]),
                                  c[..., 3] - abs(c[..., 1]) - abs(c[..., 4])),
                       c[..., 5] - abs(c[..., 2]) - abs(c[..., 4]))
    if gersh.min() < 1.0e-6:
        m = np.stack([np.stack([c[..., 0], c[..., 1], c[..., 2]], -1),
                      np.stack([c[..., 1], c[..., 3], c[..., 4]], -1),
                      np.stack([c[..., 2], c[..., 4], c[..., 5]], -1)], -2)
        lam_, vec = np.linalg.eigh(m)
        lam_ = np.maximum(lam_, 1.0e-6)
        m = (vec * lam_[..., None, :]) @ np.swapaxes(vec, -1, -2)
        c = np.stack([m[..., 0, 0], m[..., 0, 1], m[..., 0, 2], m[..., 1, 1], m[..., 1, 2], m[..., 2, 2]], -1)
    return c


def _ghost_c(ci):
    return np.pad(ci, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap")


def _noise_velocity(seed, nx, ny, nz, amp, dx, dy, dz, walls, offset):
    n = nx * ny * nz
    g = seeded_normal(seed, 3 * n, offset=offset).reshape(3, nz, ny, nx) * amp
    u, v, w = g[0], g[1], g[2]
    if walls:
        return _project_walls(u, v, w, dx, dy, dz)
    return _project_periodic(u, v, w, dx, dy, dz)


# -------------------------------------------------------------- configs
def taylor_green(n=32, steps=200, **kw):
    """Config 1: Newtonian SMAC (beta = 1 takes the fenep.f90:2479 shortcut,
    identical to smac.f90's method 1), 2 pi periodic box, Re = 100,
    dt = 0.005, u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 at the
    staggered positions U(i) at x = i dX, y = (j-1/2) dY."""
    L = 2 * math.pi
    p = default_params(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, dt=0.005, re=100.0, beta=1.0,
                       wi=2.0, lp=10.0, method=METHOD_FFT, **kw)
    d = L / n
    i = np.arange(1, n + 1)
    xf, xc = i * d, (i - 0.5) * d
    z3 = lambda a, b, c: (a[None, None, :], b[None, :, None], c[:, None, None])
    X, Y, Z = z3(xf, xc, xc)
    U = np.sin(X) * np.cos(Y) * np.cos(Z)
    X, Y, Z = z3(xc, xf, xc)
    V = -np.cos(X) * np.sin(Y) * np.cos(Z)
    W = np.zeros((n, n, n))
    c = np.zeros((n, n, n, 6))
    st = State(pad_periodic(U), pad_periodic(V), pad_periodic(W), ghosted_zeros(n, n, n), _ghost_c(c))
    return Case("cfg1_taylor_green", p, st, nsteps=steps)


def kolmogorov(n=64, nz=None, steps=1000, seed=2, lz=None, precision=0, **kw):
    """Config 2: FENE-P, periodic 2 pi box, Kolmogorov forcing
    f_x = 0.1 sin(y), Re = 100, Wi = 2, beta = 0.9, L^2 = 100, dt = 0.002."""
    nz = n if nz is None else nz
    L = 2 * math.pi
    d = L / n
    lz = nz * d if lz is None else lz
    p = default_params(nx=n, ny=n, nz=nz, lx=L, ly=L, lz=lz, dt=0.002, re=100.0, beta=0.9,
                       wi=2.0, lp=10.0, method=METHOD_FFT, precision=precision, seed=seed, **kw)
    dz = lz / nz
    u, v, w = _noise_velocity(seed, n, n, nz, 1.0e-3, d, d, dz, False, offset=0)
    c = _spd_perturbed_c(seed, n, n, nz, offset=3 * n * n * nz + 8)
    yc = (np.arange(1, n + 1) - 0.5) * d
    fx = np.broadcast_to(0.1 * np.sin(yc)[None, :, None], (nz, n, n)).copy()
    st = State(pad_periodic(u), pad_periodic(v), pad_periodic(w), ghosted_zeros(n, n, nz), _ghost_c(c))
    return Case("cfg2_kolmogorov", p, st, force=(fx, np.zeros_like(fx), np.zeros_like(fx)), nsteps=steps)


def _channel(name, nx, ny, nz, lx, ly, lz, re, wi, beta, lp, dt, centres, diameter, seed,
             poisson_variant, precision, steps, **kw):
    nu = 1.0 / re
    G = 12.0 * nu * 1.0 / ly ** 2       # plane Poiseuille gradient for mean U = 1
    p = default_params(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, dt=dt, re=re, beta=beta, wi=wi,
                       lp=lp, method=METHOD_IBM, walls_y=1, poisson_variant=poisson_variant,
                       precision=precision, body_force=(G, 0.0, 0.0), seed=seed, **kw)
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    u, v, w = _noise_velocity(seed, nx, ny, nz, 1.0e-3, dx, dy, dz, True, offset=0)
    u = u + 1.0
    c = _spd_perturbed_c(seed, nx, ny, nz, offset=3 * nx * ny * nz + 8)
    st = State(pad_periodic(u), pad_periodic(v), pad_periodic(w), ghosted_zeros(nx, ny, nz), _ghost_c(c))
    mk = cylinder_markers(centres, diameter, dx, dz, nz)
    return Case(name, p, st, markers=mk, nsteps=steps)


def channel_cylinder(nx=128, ny=64, nz=64, poisson_variant=POISSON_FFT, precision=0, steps=2000,
                     seed=3, dt=0.002, **kw):
    """Config 3: FENE-P + IBM, Lx = 8, Ly = Lz = 4, walls in y, periodic in x
    and z, one cylinder D = 0.5 at (2, 2), Re_D = 100 (re = 200), Wi = 1,
    beta = 0.8, L^2 = 100, mean flow driven by a constant pressure gradient."""
    s = nx / 128.0
    return _channel("cfg3_channel_cylinder", nx, ny, nz, 8.0 * s, 4.0 * ny / 64.0, 4.0 * nz / 64.0,
                    200.0, 1.0, 0.8, 10.0, dt, [(2.0 * s, 2.0 * ny / 64.0)], 0.5, seed,
                    poisson_variant, precision, steps, **kw)


def throughput(nx=256, ny=128, nz=128, precision=0, steps=1000, seed=4, ncyl=8, **kw):
    """Config 4: FENE-P + 8 cylinders, Lx = 16, Ly = Lz = 8 (dX = 1/16),
    walls in y, Re_D = 200 (re = 400), Wi = 4, beta = 0.9, L^2 = 400,
    dt = 0.001, seeded non-overlapping cylinders of D = 0.5 with >= 2D
    clearance from each other and from the walls."""
    dx = 1.0 / 16.0
    lx, ly, lz = nx * dx, ny * dx, nz * dx
    centres = seeded_cylinder_centres(seed, ncyl, 0.5, lx, ly)
    return _channel("cfg4_throughput", nx, ny, nz, lx, ly, lz, 400.0, 4.0, 0.9, 20.0, 0.001,
                    centres, 0.5, seed, POISSON_FFT, precision, steps, **kw)


def adjoint_case(nx=64, ny=64, nz=32, steps=400, seed=5, **kw):
    """Config 5: config-2 physics on 64x64x32 (Lz = pi keeps dX = dZ), one
    cylinder of D = 0.5 for the drag objective at x = pi, y = pi/2, where the
    Kolmogorov forcing 0.1 sin y and the flow it drives are largest (at
    y = pi the flow and the drag vanish)."""
    L = 2 * math.pi
    d = L / nx
    case = kolmogorov(n=nx, nz=nz, steps=steps, seed=seed, lz=nz * d, **kw)
    p = case.params
    p.method = METHOD_IBM
    case.markers = cylinder_markers([(L / 2, L / 4)], 0.5, d, d, nz)
    case.name = "cfg5_adjoint"
    return case


CONFIGS = {1: taylor_green, 2: kolmogorov, 3: channel_cylinder, 4: throughput, 5: adjoint_case}
