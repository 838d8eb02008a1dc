"""ctypes mirror of ``fp_params`` / ``fp_diag`` from include/fenep_gpu.h.

Parameter names follow the reference's module constants (fenep.f90:4-49).
"""
import ctypes as C

METHOD_EXPLICIT, METHOD_FFT, METHOD_IBM = 0, 1, 2
POISSON_FFT, POISSON_SOR = 0, 1
PREC_F64, PREC_MIXED = 0, 1


class FpParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("dt", C.c_double), ("re", C.c_double), ("beta", C.c_double),
        ("wi", C.c_double), ("lp", C.c_double),
        ("method", C.c_int32), ("walls_y", C.c_int32),
        ("poisson_variant", C.c_int32), ("precision", C.c_int32),
        ("newton_itrmax", C.c_int32), ("newton_eps", C.c_double),
        ("newton_dc", C.c_double),
        ("sor_itrmax", C.c_int32), ("sor_eps", C.c_double),
        ("sor_omega", C.c_double),
        ("body_force", C.c_double * 3),
        ("n_dc_max", C.c_int32), ("dc_tol", C.c_double),
        ("helm_sweeps", C.c_int32), ("reserved0", C.c_int32),
        ("ibm_weight", C.c_double), ("seed", C.c_uint64),
        ("reserved", C.c_int32 * 2),
    ]

    def copy(self):
        out = FpParams()
        C.pointer(out)[0] = self
        return out

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name == "reserved":
                continue
            d[name] = list(v) if name == "body_force" else v
        return d


class FpDiag(C.Structure):
    _fields_ = [
        ("ke", C.c_double), ("cfl", C.c_double),
        ("trc_min", C.c_double), ("trc_max", C.c_double),
        ("trc_mean", C.c_double), ("trc_std", C.c_double),
        ("div_max", C.c_double), ("drag", C.c_double), ("lift", C.c_double),
        ("newton_itr_mean", C.c_double), ("sor_itr", C.c_double),
        ("n_dc", C.c_double), ("dc_residual", C.c_double), ("sel_guarded", C.c_double),
        ("newton_unconverged", C.c_double),
        ("nan_flag", C.c_int32), ("step", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_params(**kw):
    """Reference defaults (fenep.f90:4-49 with method=1) overridden by kw."""
    import math
    p = FpParams()
    p.nx = p.ny = p.nz = 32
    p.lx = p.ly = p.lz = 2.0 * math.pi
    p.dt, p.re, p.beta, p.wi, p.lp = 0.001, 2.0, 0.9, 2.0, 100.0
    p.method, p.walls_y, p.poisson_variant, p.precision = METHOD_FFT, 0, POISSON_FFT, PREC_F64
    p.newton_itrmax, p.newton_eps, p.newton_dc = 10000, 1.0e-6, 1.0e-5
    p.sor_itrmax, p.sor_eps, p.sor_omega = 10000, 1.0e-5, 0.0
    p.body_force[0] = p.body_force[1] = p.body_force[2] = 0.0
    p.n_dc_max, p.dc_tol = 2, 1.0e-5
    for k, v in kw.items():
        if k == "body_force":
            for i in range(3):
                p.body_force[i] = float(v[i])
        elif not hasattr(p, k):
            raise KeyError(k)
        else:
            setattr(p, k, v)
    return p
