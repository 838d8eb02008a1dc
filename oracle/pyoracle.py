"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from fortran8_amd.params import FpParams, FpDiag
from fortran8_amd.fields import State

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = C.POINTER(C.c_double)


def build(force=False):
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-C", _HERE, "-s"] + (["-B"] if force else []))
    return so


def lib():
    global _LIB
    if _LIB is None:
        # FENEP_ORACLE_LIB: a library built from a modified copy of the
        # restatement (tests/test_oracle_mutations.py)
        so = os.environ.get("FENEP_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(FpParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in ["orc_set_state"]:
            getattr(L, name).argtypes = [C.c_void_p] + [_dp] * 6 + [C.c_int]
        L.orc_get_state.argtypes = [C.c_void_p] + [_dp] * 6
        L.orc_get_phi.argtypes = [C.c_void_p, _dp]
        L.orc_set_body_force.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_set_markers.argtypes = [C.c_void_p, C.c_int, C.c_int] + [_dp] * 6
        L.orc_get_marker_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_diagnostics.argtypes = [C.c_void_p, C.POINTER(FpDiag)]
        L.orc_time_stats.argtypes = [C.c_void_p, _dp]
        L.orc_selection_counts.argtypes = [C.c_void_p, _dp]
        L.orc_eval_operator.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_solve.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_sor.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_ibm_apply.argtypes = [C.c_void_p] + [_dp] * 4
        L.orc_vieta.restype = C.c_double
        L.orc_vieta.argtypes = [_dp]
        L.orc_select.argtypes = [_dp] * 5
        L.orc_gauss6.argtypes = [_dp] * 3
        L.orc_newton_cell.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_int, _dp]
        L.orc_roma_weight.restype = C.c_double
        L.orc_roma_weight.argtypes = [C.c_double]
        L.orc_tdma.argtypes = [C.c_int] + [_dp] * 7
        L.orc_fft.argtypes = [C.c_int, _dp, C.c_int]
        _LIB = L
    return _LIB


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Oracle:
    """CPU restatement of the step; same parameters and layouts as the GPU
    C-ABI."""

    def __init__(self, params):
        self.p = params.copy()
        self.h = lib().orc_create(C.byref(self.p))
        if not self.h:
            raise ValueError("oracle rejected parameters")
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, st):
        hist = None if st.hist is None else np.ascontiguousarray(st.hist, dtype=np.float64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.U, st.V, st.W, st.P, st.C)]
        lib().orc_set_state(self.h, *[ptr(a) for a in arrs], ptr(hist), int(st.has_history))

    def get_state(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        g = lambda: np.zeros((nz + 2, ny + 2, nx + 2))
        U, V, W, P = g(), g(), g(), g()
        Cc = np.zeros((nz + 2, ny + 2, nx + 2, 6))
        hist = np.zeros((9, nz, ny, nx))
        lib().orc_get_state(self.h, ptr(U), ptr(V), ptr(W), ptr(P), ptr(Cc), ptr(hist))
        return State(U, V, W, P, Cc, hist, 1)

    def get_phi(self):
        a = np.zeros((self.nz + 2, self.ny + 2, self.nx + 2))
        lib().orc_get_phi(self.h, ptr(a))
        return a

    def set_body_force(self, f):
        if f is None:
            lib().orc_set_body_force(self.h, None, None, None)
            return
        self._f = [np.ascontiguousarray(a, dtype=np.float64) for a in f]
        lib().orc_set_body_force(self.h, *[ptr(a) for a in self._f])

    def set_markers(self, mk):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (mk.xc, mk.yc, mk.zc, mk.uc, mk.vc, mk.wc)]
        lib().orc_set_markers(self.h, mk.nc, mk.nl, *[ptr(a) for a in arrs])
        self._mshape = mk.xc.shape

    def marker_forces(self):
        out = [np.zeros(self._mshape) for _ in range(3)]
        lib().orc_get_marker_forces(self.h, *[ptr(a) for a in out])
        return out

    def ibm_apply(self, up3):
        """ibm_Helmholtz on a ghosted (3, nz+2, ny+2, nx+2) velocity; returns
        (spread force (3, nz, ny, nx), marker forces, marker velocities)."""
        up3 = np.ascontiguousarray(up3, dtype=np.float64)
        fint = np.zeros((3, self.nz, self.ny, self.nx))
        fm = np.zeros((3,) + self._mshape)
        ub = np.zeros((3,) + self._mshape)
        rc = lib().orc_ibm_apply(self.h, ptr(up3), ptr(fint), ptr(fm), ptr(ub))
        if rc:
            raise RuntimeError("ibm_apply rc=%d" % rc)
        return fint, fm, ub

    def step(self, n=1):
        rc = lib().orc_step(self.h, int(n))
        if rc:
            raise RuntimeError("oracle step failed rc=%d" % rc)

    def diagnostics(self):
        d = FpDiag()
        lib().orc_diagnostics(self.h, C.byref(d))
        return d.as_dict()

    def time_stats(self):
        """all_time_dif statistics (13 values, fenep_mpi.f90:2342-2344 order)."""
        out = np.zeros(13)
        lib().orc_time_stats(self.h, ptr(out))
        return out

    def energy_spectrum(self, nbins=None):
        """Shell spectrum of the regular-grid velocity (fft_energy_dif,
        fenep_mpi.f90:2566-2665, no subtracted mean), with numpy's FFT as the
        transform: |u^/N|^2 summed into shells nint(|k|), half weight on the
        kx = 0 and Nyquist planes of the half spectrum."""
        st = self.get_state()
        nz, ny, nx = st.U.shape[0] - 2, st.U.shape[1] - 2, st.U.shape[2] - 2
        N = nx * ny * nz
        I = (slice(1, -1), slice(1, -1), slice(1, -1))
        ureg = (st.U[1:-1, 1:-1, 1:-1] + st.U[1:-1, 1:-1, :-2]) / 2.0
        vreg = (st.V[1:-1, 1:-1, 1:-1] + st.V[1:-1, :-2, 1:-1]) / 2.0
        wreg = (st.W[1:-1, 1:-1, 1:-1] + st.W[:-2, 1:-1, 1:-1]) / 2.0
        e = sum(np.abs(np.fft.rfftn(a) / N) ** 2 for a in (ureg, vreg, wreg))
        kx = np.arange(nx // 2 + 1, dtype=float)[None, None, :]
        ky = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny).astype(float)[None, :, None]
        kz = np.where(np.arange(nz) <= nz // 2, np.arange(nz), np.arange(nz) - nz).astype(float)[:, None, None]
        idx = np.floor(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) + 0.5).astype(int)
        w = np.ones(nx // 2 + 1)
        w[0] = w[-1] = 0.5
        nb = nx if nbins is None else int(nbins)
        E = np.bincount(idx.ravel(), weights=(e * w[None, None, :]).ravel(), minlength=idx.max() + 1)
        return E[:nb] if len(E) >= nb else np.pad(E, (0, nb - len(E)))

    def selection_counts(self):
        c = np.zeros(4)
        lib().orc_selection_counts(self.h, ptr(c))
        return c

    def eval_operator(self, which):
        out = [np.zeros((self.nz, self.ny, self.nx)) for _ in range(3)]
        lib().orc_eval_operator(self.h, which, *[ptr(a) for a in out])
        return out

    def solve(self, kind, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        x = np.zeros_like(q)
        lib().orc_solve(self.h, kind, ptr(q), ptr(x))
        return x

    def sor(self, q, phi0=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        phi = np.zeros((self.nz + 2, self.ny + 2, self.nx + 2)) if phi0 is None else phi0.copy()
        it = lib().orc_sor(self.h, ptr(q), ptr(phi))
        return phi, it


def load_case(case):
    o = Oracle(case.params)
    o.set_state(case.state)
    if case.force is not None:
        o.set_body_force(case.force)
    if case.markers is not None:
        o.set_markers(case.markers)
    return o


# ---- unit-level helpers
def vieta(a6):
    a = np.ascontiguousarray(a6, dtype=np.float64)
    return lib().orc_vieta(ptr(a))


def select(cm, c, cp):
    cm, c, cp = [np.ascontiguousarray(x, dtype=np.float64) for x in (cm, c, cp)]
    po, no = np.zeros(6), np.zeros(6)
    idx = lib().orc_select(ptr(cm), ptr(c), ptr(cp), ptr(po), ptr(no))
    return idx, po, no


def gauss6(A, b):
    A = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(6)
    lib().orc_gauss6(ptr(A), ptr(b), ptr(x))
    return x


def newton_cell(A, b, x0, dt, wi, lp, dc=1e-5, eps=1e-6, itrmax=10000):
    A = np.asarray(A, dtype=np.float64).ravel(order="F").copy()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    x = np.zeros(6)
    it = lib().orc_newton_cell(ptr(A), ptr(b), ptr(x0), dt, wi, lp, dc, eps, itrmax, ptr(x))
    return x, it


def roma(r):
    return lib().orc_roma_weight(float(r))


def tdma(a, b, c, d):
    n = len(a)
    a, b, c = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c)]
    d = np.asarray(d, dtype=np.complex128)
    dr, di = np.ascontiguousarray(d.real), np.ascontiguousarray(d.imag)
    xr, xi = np.zeros(n), np.zeros(n)
    lib().orc_tdma(n, ptr(a), ptr(b), ptr(c), ptr(dr), ptr(di), ptr(xr), ptr(xi))
    return xr + 1j * xi


def fft(x, inverse=False):
    x = np.asarray(x, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel())
    lib().orc_fft(len(x), ptr(buf), 1 if inverse else 0)
    return buf[0::2] + 1j * buf[1::2]
