"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from fortran8_amd.params import FpParams, FpDiag
from fortran8_amd.fields import State, interior, interior_c

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

    # ---- statistics with a subtracted mean and per-column maps (numpy)
    def mean_accumulate(self):
        """fenep_mpi.f90:2969-2971: U_ave += U (interior)."""
        st = self.get_state()
        cur = [interior(getattr(st, n)).copy() for n in "UVW"]
        if getattr(self, "_mean_acc", None) is None:
            self._mean_acc, self._mean_cnt = cur, 1
        else:
            self._mean_acc = [a + b for a, b in zip(self._mean_acc, cur)]
            self._mean_cnt += 1

    def mean_apply(self, mode):
        """0: no mean; 1: accumulated mean / count, averaged over z (ave_mean,
        fenep_mpi.f90:2010-2033, 3000-3003); 2: not averaged over z."""
        if mode == 0:
            self._mean = None
        else:
            m = [a / self._mean_cnt for a in self._mean_acc]
            if mode == 1:
                m = [np.broadcast_to(a.sum(axis=0) / a.shape[0], a.shape).copy() for a in m]
            self._mean = m
        self._mean_acc = None

    def _fluct(self):
        st = self.get_state()
        walls = bool(self.p.walls_y)
        mean = getattr(self, "_mean", None)
        out = []
        for a, nm in enumerate("UVW"):
            u = getattr(st, nm)
            if mean is not None:
                u = u - ghost_fill(mean[a], "UVW"[a], walls)
            out.append(u)
        return st, out

    def time_stats_np(self, step=None):
        """all_time_dif (fenep_mpi.f90:2200-2348) in numpy, of the velocity
        less the applied mean."""
        st, (u, v, w) = self._fluct()
        t = stat_terms(st, u, v, w, self.p)
        N = float(self.nx * self.ny * self.nz)
        sums = [x.sum() / N for x in t]
        return stat_record(sums, self.p, self._step(step), var=((t[9] - sums[9]) ** 2).sum() / N,
                           kst=sums[7] / 2.0)

    def time_stats_map(self, centres=(), diameter=0.0, step=None):
        """all_time_dif_map (fenep_mpi.f90:2350-2564): the per-(i, j)
        statistics of the z columns, the 18-column map record (j outer, i
        inner) and the plane averages over the unmasked columns."""
        st, (u, v, w) = self._fluct()
        t = stat_terms(st, u, v, w, self.p)
        nz = float(self.nz)
        col = [x.sum(axis=0) / nz for x in t]           # (ny, nx)
        var = ((t[9] - col[9][None]) ** 2).sum(axis=0) / nz
        nu = 1.0 / self.p.re
        L2 = self.p.lp ** 2
        urms, vrms, wrms = col[0], col[1], col[2]
        all_rms = (urms + vrms + wrms) / 3.0
        all_grad = (col[3] + col[4] + col[5]) / 3.0
        with np.errstate(divide="ignore", invalid="ignore"):
            eps = nu / 2.0 * col[6]
            Lk, Tk, Uk = (nu ** 3 / eps) ** 0.25, (nu / eps) ** 0.5, (nu * eps) ** 0.25
            lam = np.sqrt(all_rms / all_grad)
            relam = np.sqrt(all_rms) * lam / nu
        K = (urms + vrms + wrms) / 2.0
        Ep = 0.5 * (1.0 - self.p.beta) / self.p.re / self.p.wi * (L2 - 3.0) * col[8]
        mask = plane_mask(self.nx, self.ny, self.p.lx / self.nx, self.p.ly / self.ny, centres, diameter)
        dx, dy = self.p.lx / self.nx, self.p.ly / self.ny
        X, Y = np.meshgrid((np.arange(self.nx) + 0.5) * dx, (np.arange(self.ny) + 0.5) * dy)
        vals = [K, lam, relam, eps, Lk, Tk, Uk, Ep, col[9], np.sqrt(var)]
        rows = np.zeros((self.ny, self.nx, 18))
        rows[..., 0], rows[..., 1] = X, Y
        for q, a in enumerate(vals):
            rows[..., 2 + q] = a
        rows[..., 12] = mask
        ms = mask.sum()
        with np.errstate(invalid="ignore"):
            avg = [(a * mask).sum() / ms for a in vals]
        rec = np.array([self._step(step), self._step(step) * self.p.dt, avg[0], 0.0, avg[1], avg[2], avg[3],
                        avg[4], avg[5], avg[6], avg[7], avg[8], avg[9]])
        return rows.reshape(-1, 18), rec

    def energy_spectrum_z(self, centres=(), diameter=0.0):
        """fft_energy_dif_map (fenep_mpi.f90:2670-2797): numpy's FFT along z
        of each column's regular-grid velocity / NZ (fft_z,
        fenep_mpi.f90:1442-1463); energy per |kz| = 0..NZ/2, half weight at
        kz = 0 and NZ/2; plane average over the unmasked columns and the
        column (1, 1)."""
        st, (u, v, w) = self._fluct()
        ur = (u[1:-1, 1:-1, 1:-1] + u[1:-1, 1:-1, :-2]) / 2.0
        vr = (v[1:-1, 1:-1, 1:-1] + v[1:-1, :-2, 1:-1]) / 2.0
        wr = (w[1:-1, 1:-1, 1:-1] + w[:-2, 1:-1, 1:-1]) / 2.0
        nz = self.nz
        e = sum(np.abs(np.fft.fft(a, axis=0) / nz) ** 2 for a in (ur, vr, wr))   # (nz, ny, nx)
        E = np.zeros_like(e)
        for k in range(nz // 2 + 1):
            E[k] = e[k] / 2.0 if k in (0, nz // 2) else e[k] / 2.0 * 2.0
        mask = plane_mask(self.nx, self.ny, self.p.lx / self.nx, self.p.ly / self.ny, centres, diameter)
        em = (E * mask[None]).sum(axis=(1, 2)) / mask.sum()
        return em, E[:, 0, 0].copy()

    def _step(self, step):
        return float(self.diagnostics()["step"] if step is None else step)

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


def ghost_fill(a, kind, walls):
    """Ghosted copy of an interior (nz, ny, nx) array by the boundary rules:
    periodic in x and z; in y periodic, or the wall rules U_kabe/V_kabe/
    W_kabe (fenep_2d.f90:1098-1143): tangential ghost = -interior, V zero on
    the wall face and ghosts."""
    g = np.pad(a, 1, mode="wrap")
    if walls:
        if kind in "UW":
            g[:, 0, :] = -g[:, 1, :]
            g[:, -1, :] = -g[:, -2, :]
        elif kind == "V":
            g[:, 0, :] = 0.0
            g[:, -2, :] = 0.0
            g[:, -1, :] = 0.0
    return g


def stat_terms(st, U, V, W, p):
    """Per-cell terms of all_time_dif (fenep_mpi.f90:2229-2330) of the
    ghosted fluctuation velocity U, V, W: regular-grid squares, staggered
    gradient squares, S:S of the 4-point gradient, |u|^2 of the raw staggered
    velocity, log f(C), tr C; each (nz, ny, nx)."""
    dx, dy, dz = p.lx / p.nx, p.ly / p.ny, p.lz / p.nz
    sl = lambda a, di, dj, dk: a[1 + dk:a.shape[0] - 1 + dk, 1 + dj:a.shape[1] - 1 + dj, 1 + di:a.shape[2] - 1 + di]
    ur = (sl(U, 0, 0, 0) + sl(U, -1, 0, 0)) / 2.0
    vr = (sl(V, 0, 0, 0) + sl(V, 0, -1, 0)) / 2.0
    wr = (sl(W, 0, 0, 0) + sl(W, 0, 0, -1)) / 2.0
    D = np.zeros((3, 3) + ur.shape)
    D[0, 0] = (sl(U, 0, 0, 0) - sl(U, -1, 0, 0)) / dx
    D[0, 1] = (sl(U, 0, 1, 0) - sl(U, 0, -1, 0) + sl(U, -1, 1, 0) - sl(U, -1, -1, 0)) / (4 * dy)
    D[0, 2] = (sl(U, 0, 0, 1) - sl(U, 0, 0, -1) + sl(U, -1, 0, 1) - sl(U, -1, 0, -1)) / (4 * dz)
    D[1, 0] = (sl(V, 1, 0, 0) - sl(V, -1, 0, 0) + sl(V, 1, -1, 0) - sl(V, -1, -1, 0)) / (4 * dx)
    D[1, 1] = (sl(V, 0, 0, 0) - sl(V, 0, -1, 0)) / dy
    D[1, 2] = (sl(V, 0, 0, 1) - sl(V, 0, 0, -1) + sl(V, 0, -1, 1) - sl(V, 0, -1, -1)) / (4 * dz)
    D[2, 0] = (sl(W, 1, 0, 0) - sl(W, -1, 0, 0) + sl(W, 1, 0, -1) - sl(W, -1, 0, -1)) / (4 * dx)
    D[2, 1] = (sl(W, 0, 1, 0) - sl(W, 0, -1, 0) + sl(W, 0, 1, -1) - sl(W, 0, -1, -1)) / (4 * dy)
    D[2, 2] = (sl(W, 0, 0, 0) - sl(W, 0, 0, -1)) / dz
    S = D + np.swapaxes(D, 0, 1)
    raw = sum(interior(getattr(st, n)) ** 2 for n in "UVW")
    C = interior_c(st.C)
    tr = C[..., 0] + C[..., 3] + C[..., 5]
    L2 = p.lp ** 2
    return [ur ** 2, vr ** 2, wr ** 2, D[0, 0] ** 2, D[1, 1] ** 2, D[2, 2] ** 2, (S ** 2).sum(axis=(0, 1)), raw,
            np.log((L2 - 3.0) / (L2 - tr)), tr]


def stat_record(m, p, step, var, kst):
    """The 13 values of an all_time_dif record from the plane/volume means m
    of the stat_terms (fenep_mpi.f90:2311-2344)."""
    nu = 1.0 / p.re
    all_rms = (m[0] + m[1] + m[2]) / 3.0
    all_grad = (m[3] + m[4] + m[5]) / 3.0
    eps = nu / 2.0 * m[6]
    with np.errstate(divide="ignore", invalid="ignore"):
        return _record(m, p, step, var, kst, nu, all_rms, all_grad, eps)


def _record(m, p, step, var, kst, nu, all_rms, all_grad, eps):
    lam = np.sqrt(all_rms / all_grad)
    return np.array([step, step * p.dt, (m[0] + m[1] + m[2]) / 2.0, kst, lam, np.sqrt(all_rms) * lam / nu, eps,
                     (nu ** 3 / eps) ** 0.25, (nu / eps) ** 0.5, (nu * eps) ** 0.25,
                     0.5 * (1.0 - p.beta) / p.re / p.wi * (p.lp ** 2 - 3.0) * m[8], m[9], np.sqrt(var)])


def plane_mask(nx, ny, dx, dy, centres, diameter):
    """1 outside 1.5 radii of every cylinder centre, 0 inside
    (fenep_mpi.f90:2530-2537 with the centres given)."""
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy)
    mask = np.ones((ny, nx))
    r = diameter * 1.5 / 2.0
    for (xc, yc) in np.asarray(centres, dtype=float).reshape(-1, 2):
        mask[(X - xc) ** 2 + (Y - yc) ** 2 < r * r] = 0.0
    return mask


_CS_LIB = None


def cs_lib():
    """The complex-step build of the restatement (liboracle_cs.so)."""
    global _CS_LIB
    if _CS_LIB is None:
        so = os.path.join(_HERE, "liboracle_cs.so")
        if not os.path.exists(so):
            build(force=True)
        L = C.CDLL(so)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(FpParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p] + [_dp] * 6 + [C.c_int]
        L.orc_get_state.argtypes = [C.c_void_p] + [_dp] * 6
        L.orc_set_body_force.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_set_markers.argtypes = [C.c_void_p, C.c_int, C.c_int] + [_dp] * 6
        L.orc_get_marker_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_cs_seed_param.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_cs_seed_force.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_cs_drag.argtypes = [C.c_void_p, _dp]
        _CS_LIB = L
    return _CS_LIB


class ComplexStep:
    """The restatement evaluated in complex-step arithmetic (oracle/cstep.h):
    the real parts are the FP64 oracle's bit for bit, the imaginary parts h
    times the derivative along one seeded direction."""

    def __init__(self, case):
        self.case = case
        self.h = cs_lib().orc_create(C.byref(case.params))
        if not self.h:
            raise ValueError("oracle rejected parameters")
        L = cs_lib()
        st = case.state
        hist = None if st.hist is None else np.ascontiguousarray(st.hist, dtype=np.float64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.U, st.V, st.W, st.P, st.C)]
        L.orc_set_state(self.h, *[ptr(a) for a in arrs], ptr(hist), int(st.has_history))
        if case.force is not None:
            self._f = [np.ascontiguousarray(a, dtype=np.float64) for a in case.force]
            L.orc_set_body_force(self.h, *[ptr(a) for a in self._f])
        if case.markers is not None:
            mk = case.markers
            self._mk = [np.ascontiguousarray(a, dtype=np.float64) for a in (mk.xc, mk.yc, mk.zc, mk.uc, mk.vc, mk.wc)]
            L.orc_set_markers(self.h, mk.nc, mk.nl, *[ptr(a) for a in self._mk])

    def seed_param(self, which, h):
        """which: 0 Wi, 1 beta, 2 L^2."""
        cs_lib().orc_cs_seed_param(self.h, int(which), float(h))

    def seed_force(self, dF, h):
        d = np.ascontiguousarray(dF, dtype=np.float64)
        if cs_lib().orc_cs_seed_force(self.h, ptr(d), float(h)):
            raise RuntimeError("the case has no body-force field to seed")

    def drag_series(self, nsteps):
        """Steps nsteps times; returns (drag, imaginary part of drag) per step."""
        L = cs_lib()
        re, im = np.zeros(nsteps), np.zeros(nsteps)
        out = np.zeros(2)
        for s in range(nsteps):
            L.orc_step(self.h, 1)
            L.orc_cs_drag(self.h, ptr(out))
            re[s], im[s] = out
        return re, im

    def get_state(self):
        nx, ny, nz = self.case.params.nx, self.case.params.ny, self.case.params.nz
        g = lambda: np.zeros((nz + 2, ny + 2, nx + 2))
        U, V, W, P = g(), g(), g(), g()
        Cc = np.zeros((nz + 2, ny + 2, nx + 2, 6))
        hist = np.zeros((9, nz, ny, nx))
        cs_lib().orc_get_state(self.h, ptr(U), ptr(V), ptr(W), ptr(P), ptr(Cc), ptr(hist))
        return State(U, V, W, P, Cc, hist, 1)

    def close(self):
        if self.h:
            cs_lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


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
