"""Host-side state: layouts, seeded inputs, cylinder markers.

Layouts are the reference's (Fortran column-major, x fastest):
  ghosted scalar  (0:NX+1, 0:NY+1, 0:NZ+1)  -> numpy shape (NZ+2, NY+2, NX+2)
  tensor C        (6, 0:NX+1, 0:NY+1, 0:NZ+1) -> numpy shape (NZ+2, NY+2, NX+2, 6)
  histories       (1:NX, 1:NY, 1:NZ)        -> numpy shape (NZ, NY, NX)
"""
from dataclasses import dataclass, field
import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def seeded_uniform(seed, n, offset=0):
    """SplitMix64 of (seed, offset+i+1), top 53 bits -> [0,1).

    Same stream as ``fp_seeded_uniform`` in the C-ABI library."""
    idx = np.arange(offset + 1, offset + 1 + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def seeded_normal(seed, n, offset=0):
    """Box-Muller on pairs of the uniform stream."""
    m = (n + 1) // 2
    u = seeded_uniform(seed, 2 * m, offset)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


@dataclass
class State:
    """Full step state in the reference layout."""
    U: np.ndarray
    V: np.ndarray
    W: np.ndarray
    P: np.ndarray
    C: np.ndarray                  # (nz+2, ny+2, nx+2, 6)
    hist: np.ndarray = None        # (9, nz, ny, nx): A0x..z, T0x..z, B0x..z
    has_history: int = 0

    def copy(self):
        return State(self.U.copy(), self.V.copy(), self.W.copy(), self.P.copy(),
                     self.C.copy(), None if self.hist is None else self.hist.copy(),
                     self.has_history)


def ghosted_zeros(nx, ny, nz):
    return np.zeros((nz + 2, ny + 2, nx + 2), dtype=np.float64)


def pad_periodic(a):
    """interior (nz,ny,nx) -> ghosted with periodic copies (PBM)."""
    return np.pad(a, 1, mode="wrap")


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def interior_c(c):
    return c[1:-1, 1:-1, 1:-1, :]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


# ---------------------------------------------------------------- markers
@dataclass
class Markers:
    nc: int
    nl: int
    xc: np.ndarray   # (nz, nl, nc) i.e. Fortran (NC, NL, NZ)
    yc: np.ndarray
    zc: np.ndarray
    uc: np.ndarray
    vc: np.ndarray
    wc: np.ndarray
    centers: list = field(default_factory=list)
    diameter: float = 0.0


def cylinder_markers(centers, diameter, dx, dz, nz, interior_shells=True):
    """Marker layout of ibm_init (fenep.f90:1769-1900) for stationary
    z-spanning cylinders at the given (x, y) centres.

    interior_shells=True is ibm_type 2x: shells of NL_shell(i) =
    nint(pi (D - 2 i dX)/dX) points, i = 0..int(D/(2 dX)), a single centre
    point when the last shell is empty.  False is ibm_type 1x: NL =
    nint(pi D/dX) surface points.  Plane n has Zc = (n - 0.25) dZ."""
    nc = len(centers)
    if interior_shells:
        nl_lap = int(diameter / (2 * dx))
        shells = [int(np.rint(math.pi * (diameter - 2 * i * dx) / dx)) for i in range(nl_lap + 1)]
        if shells[-1] == 0:
            shells[-1] = 1
        radii = [(diameter - 2 * i * dx) / 2 for i in range(nl_lap + 1)]
    else:
        shells = [int(np.rint(math.pi * diameter / dx))]
        radii = [diameter / 2]
    nl = sum(shells)
    xs = np.zeros((nl, nc))
    ys = np.zeros((nl, nc))
    idx = 0
    for s, (cnt, r) in enumerate(zip(shells, radii)):
        for j in range(1, cnt + 1):
            th = 2 * math.pi / cnt * j
            for l, (cx, cy) in enumerate(centers):
                xs[idx, l] = cx + r * math.cos(th)
                ys[idx, l] = cy + r * math.sin(th)
            idx += 1
    xc = np.broadcast_to(xs, (nz, nl, nc)).copy()
    yc = np.broadcast_to(ys, (nz, nl, nc)).copy()
    zc = np.empty((nz, nl, nc))
    for n in range(1, nz + 1):
        zc[n - 1] = (n - 0.25) * dz
    z0 = np.zeros((nz, nl, nc))
    return Markers(nc, nl, xc, yc, zc, z0.copy(), z0.copy(), z0.copy(), list(centers), diameter)


def seeded_cylinder_centres(seed, count, diameter, lx, ly, clearance=2.0, wall_clearance=2.0):
    """Non-overlapping (x, y) centres with >= clearance*D between surfaces of
    neighbouring cylinders and from the y-walls (rejection sampling on the
    seeded stream)."""
    D = diameter
    out = []
    off = 0
    while len(out) < count:
        u = seeded_uniform(seed, 2, offset=off)
        off += 2
        x = u[0] * lx
        y = (wall_clearance + 0.5) * D + u[1] * (ly - 2 * (wall_clearance + 0.5) * D)
        ok = x > (clearance + 0.5) * D and x < lx - (clearance + 0.5) * D
        for (a, b) in out:
            if math.hypot(a - x, b - y) < (1.0 + clearance) * D:
                ok = False
        if ok:
            out.append((float(x), float(y)))
        if off > 200000:
            raise RuntimeError("could not place cylinders")
    return out
