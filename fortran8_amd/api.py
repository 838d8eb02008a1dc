"""Python host binding of the C-ABI in include/fenep_gpu.h.

``FenepGPU`` mirrors the reference's step interface: the state arrays and
parameters of fenep.f90's program main (fenep.f90:2424-2446), advanced by
``step`` (the loop body of fenep.f90:2477-2527) on one MI355X.  The HIP
library must be built (``fortran8_amd.build``); there is no CPU fallback.
"""
import ctypes as C
import os
import re

import numpy as np

from .params import FpParams, FpDiag
from .fields import State

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfenep_gpu.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "fenep_gpu.h")
_LIB = None
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

ERRORS = {1: "FP_E_ARG", 2: "FP_E_HIP", 3: "FP_E_STATE", 4: "FP_E_NAN", 5: "FP_E_NOMEM"}


class FenepError(RuntimeError):
    pass


def header_exports():
    """Function names declared in include/fenep_gpu.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fp_\w+)\s*\(", txt, re.M)))


def build_id():
    """Source hash compiled into the loaded library (equal to
    fortran8_amd.build.source_hash() of this tree, or lib() refuses it)."""
    return lib().fp_build_id().decode()


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FenepError("HIP library %s is not built; run fortran8_amd/build.py" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.fp_build_id.restype = C.c_char_p
        L.fp_build_id.argtypes = []
        from .build import source_hash
        built, tree = L.fp_build_id().decode(), source_hash()
        if built != tree:
            raise FenepError("%s was built from other sources (build id %s, sources %s); rebuild it with "
                             "fortran8_amd/build.py" % (LIB_PATH, built, tree))
        L.fp_default_params.argtypes = [C.POINTER(FpParams)]
        L.fp_create.argtypes = [C.POINTER(FpParams), C.POINTER(_vp)]
        L.fp_destroy.argtypes = [_vp]
        L.fp_last_error.restype = C.c_char_p
        L.fp_last_error.argtypes = [_vp]
        L.fp_set_state.argtypes = [_vp] + [_dp] * 6 + [C.c_int]
        L.fp_get_state.argtypes = [_vp] + [_dp] * 6
        L.fp_get_phi.argtypes = [_vp, _dp]
        L.fp_set_body_force.argtypes = [_vp, _dp, _dp, _dp]
        L.fp_set_markers.argtypes = [_vp, C.c_int, C.c_int] + [_dp] * 6
        L.fp_get_marker_forces.argtypes = [_vp, _dp, _dp, _dp]
        L.fp_step.argtypes = [_vp, C.c_int, C.c_int]
        L.fp_diagnostics.argtypes = [_vp, C.POINTER(FpDiag)]
        L.fp_diagnostics_static.argtypes = [_vp, C.POINTER(FpDiag)]
        L.fp_set_state_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]
        L.fp_get_state_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.fp_enable_timing.argtypes = [_vp, C.c_int]
        L.fp_kernel_times.argtypes = [_vp, _dp, C.c_int, C.c_char_p, C.c_int]
        L.fp_seeded_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, _dp]
        L.fp_set_physics.argtypes = [_vp, C.c_double, C.c_double, C.c_double]
        L.fp_time_stats.argtypes = [_vp, _dp]
        L.fp_energy_spectrum.argtypes = [_vp, _dp, C.c_int]
        L.fp_check_selection.argtypes = [_vp, _dp]
        L.fp_get_selection.argtypes = [_vp, C.c_void_p]
        L.fp_time_stats_map.argtypes = [_vp, C.c_int, _dp, C.c_double, _dp, _dp]
        L.fp_energy_spectrum_z.argtypes = [_vp, C.c_int, _dp, C.c_double, _dp, _dp]
        L.fp_mean_accumulate.argtypes = [_vp]
        L.fp_mean_apply.argtypes = [_vp, C.c_int]
        L.fp_adj_record.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp]
        L.fp_adj_backward.argtypes = [_vp, _dp, _vp, _vp, _dp, _vp]
        _LIB = L
    return _LIB


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise FenepError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def seeded_uniform(seed, n, offset=0):
    out = np.zeros(n)
    lib().fp_seeded_uniform(C.c_uint64(seed), C.c_uint64(offset), n, _ptr(out))
    return out


class FenepGPU:
    """One device context (one HIP stream, device-resident state)."""

    def __init__(self, params):
        self.p = params.copy()
        h = _vp()
        rc = lib().fp_create(C.byref(self.p), C.byref(h))
        if rc:
            raise FenepError("fp_create failed: %s" % ERRORS.get(rc, rc))
        self.h = h
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self._mshape = None

    def _check(self, rc, what):
        if rc:
            msg = lib().fp_last_error(self.h)
            raise FenepError("%s failed: %s %s" % (what, ERRORS.get(rc, rc), msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            lib().fp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, st):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.U, st.V, st.W, st.P, st.C)]
        hist = None if st.hist is None else np.ascontiguousarray(st.hist, dtype=np.float64)
        self._check(lib().fp_set_state(self.h, *[_ptr(a) for a in arrs], _ptr(hist), int(st.has_history)),
                    "fp_set_state")

    def get_state(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        g = lambda: np.zeros((nz + 2, ny + 2, nx + 2))
        U, V, W, P = g(), g(), g(), g()
        Cc = np.zeros((nz + 2, ny + 2, nx + 2, 6))
        hist = np.zeros((9, nz, ny, nx))
        self._check(lib().fp_get_state(self.h, _ptr(U), _ptr(V), _ptr(W), _ptr(P), _ptr(Cc), _ptr(hist)),
                    "fp_get_state")
        return State(U, V, W, P, Cc, hist, 1)

    def get_phi(self):
        a = np.zeros((self.nz + 2, self.ny + 2, self.nx + 2))
        self._check(lib().fp_get_phi(self.h, _ptr(a)), "fp_get_phi")
        return a

    def set_body_force(self, f):
        if f is None:
            self._check(lib().fp_set_body_force(self.h, None, None, None), "fp_set_body_force")
            return
        self._f = [np.ascontiguousarray(a, dtype=np.float64) for a in f]
        self._check(lib().fp_set_body_force(self.h, *[_ptr(a) for a in self._f]), "fp_set_body_force")

    def set_markers(self, mk):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (mk.xc, mk.yc, mk.zc, mk.uc, mk.vc, mk.wc)]
        self._check(lib().fp_set_markers(self.h, mk.nc, mk.nl, *[_ptr(a) for a in arrs]), "fp_set_markers")
        self._mshape = mk.xc.shape

    def marker_forces(self):
        out = [np.zeros(self._mshape) for _ in range(3)]
        self._check(lib().fp_get_marker_forces(self.h, *[_ptr(a) for a in out]), "fp_get_marker_forces")
        return out

    def step(self, n=1, sync=True):
        self._check(lib().fp_step(self.h, int(n), 1 if sync else 0), "fp_step")

    def diagnostics(self):
        d = FpDiag()
        rc = lib().fp_diagnostics(self.h, C.byref(d))
        if rc and rc != 4:
            self._check(rc, "fp_diagnostics")
        return d.as_dict()

    def time_stats(self):
        """all_time_dif statistics (13 values, fenep_mpi.f90:2343-2344 order)."""
        out = np.zeros(13)
        self._check(lib().fp_time_stats(self.h, _ptr(out)), "fp_time_stats")
        return out

    def energy_spectrum(self, nbins=None):
        """Shell spectrum E(k), k = 0..nbins-1 (default NX)."""
        nb = self.nx if nbins is None else int(nbins)
        out = np.zeros(nb)
        self._check(lib().fp_energy_spectrum(self.h, _ptr(out), nb), "fp_energy_spectrum")
        return out

    def time_stats_map(self, centres=(), diameter=0.0):
        """all_time_dif_map: (map (NY*NX, 18) in record order, masked 13-value
        record); centres: (x, y) pairs of the masked cylinders."""
        xy = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1))
        m = np.zeros((self.ny * self.nx, 18))
        out = np.zeros(13)
        self._check(lib().fp_time_stats_map(self.h, len(xy) // 2, _ptr(xy) if len(xy) else None, float(diameter),
                                            _ptr(m), _ptr(out)), "fp_time_stats_map")
        return m, out

    def energy_spectrum_z(self, centres=(), diameter=0.0):
        """fft_energy_dif_map of the current state: (masked plane average of
        the column z spectra, spectrum of column (1, 1)), NZ entries each."""
        xy = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1))
        em, ef = np.zeros(self.nz), np.zeros(self.nz)
        self._check(lib().fp_energy_spectrum_z(self.h, len(xy) // 2, _ptr(xy) if len(xy) else None, float(diameter),
                                               _ptr(em), _ptr(ef)), "fp_energy_spectrum_z")
        return em, ef

    def mean_accumulate(self):
        self._check(lib().fp_mean_accumulate(self.h), "fp_mean_accumulate")

    def mean_apply(self, mode):
        """0: no subtracted mean; 1: the accumulated mean averaged over z;
        2: the accumulated mean as it is."""
        self._check(lib().fp_mean_apply(self.h, int(mode)), "fp_mean_apply")

    def check_selection(self):
        """Mixed precision: (re-decided in FP64, differing from the reference's
        selection, total) over the C+- selection of the current state."""
        out = np.zeros(3)
        self._check(lib().fp_check_selection(self.h, _ptr(out)), "fp_check_selection")
        return {"fp64": int(out[0]), "differ": int(out[1]), "total": int(out[2])}

    def get_selection(self):
        """Mixed precision: the C+- selection of the current stored field,
        (nz, ny, nx) bytes; bits 2d, 2d+1 the choice along axis d (x, y, z),
        0-2 the candidate levels, 3 the zero slope."""
        out = np.zeros((self.nz, self.ny, self.nx), dtype=np.uint8)
        self._check(lib().fp_get_selection(self.h, out.ctypes.data_as(C.c_void_p)), "fp_get_selection")
        return out

    # ---- sensitivities (FP64 path, periodic grids)
    def set_physics(self, wi, beta, lp):
        self._check(lib().fp_set_physics(self.h, float(wi), float(beta), float(lp)), "fp_set_physics")

    def adj_record(self, nsteps, every, replay=False):
        """Advance nsteps steps while recording; returns the drag of each step."""
        drag = np.zeros(nsteps)
        self._check(lib().fp_adj_record(self.h, int(nsteps), int(every), 1 if replay else 0, _ptr(drag)),
                    "fp_adj_record")
        return drag

    def adj_backward(self, weights=None, bar_final=0, bar_initial=0, bar_force=0):
        """Reverse sweep of the last recording.  weights: host array over the
        recorded steps; bar_final / bar_initial / bar_force: device pointers
        (int, 0 = none).  Returns (dJ/dWi, dJ/dbeta, dJ/dL^2)."""
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        out = np.zeros(3)
        self._check(lib().fp_adj_backward(self.h, _ptr(w), bar_final or None, bar_initial or None, _ptr(out),
                                          bar_force or None), "fp_adj_backward")
        return out

    def enable_timing(self, on=True):
        lib().fp_enable_timing(self.h, int(on) if not isinstance(on, bool) else (1 if on else 0))

    def kernel_times(self):
        ms = np.zeros(64)
        names = C.create_string_buffer(4096)
        k = lib().fp_kernel_times(self.h, _ptr(ms), 64, names, 4096)
        nm = names.value.decode().split(",") if k else []
        return dict(zip(nm, ms[:k].tolist()))


def load_case(case):
    g = FenepGPU(case.params)
    g.set_state(case.state)
    if case.force is not None:
        g.set_body_force(case.force)
    if case.markers is not None:
        g.set_markers(case.markers)
    return g
