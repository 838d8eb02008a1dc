"""fp_diagnostics_static, the entry point that allocates nothing, against
fp_diagnostics through the C-ABI: two calls in a row on one GPU, the second
on a smaller grid, FP64 and mixed contexts, with and without markers."""
import ctypes as C

import pytest

from fortran8_amd import api, configs
from fortran8_amd.params import FpDiag, PREC_MIXED

pytestmark = pytest.mark.gpu


def both(g):
    a, b = FpDiag(), FpDiag()
    ra = api.lib().fp_diagnostics(g.h, C.byref(a))
    rb = api.lib().fp_diagnostics_static(g.h, C.byref(b))
    return ra, a.as_dict(), rb, b.as_dict()


def test_static_diagnostics_equal_the_allocating_ones_on_two_calls():
    mixed = configs.channel_cylinder(nx=64, ny=32, nz=16)
    mixed.params.precision = PREC_MIXED
    cases = [("channel 64x32x16 FP64", configs.channel_cylinder(nx=64, ny=32, nz=16), 3),
             ("kolmogorov 16^3 FP64 (smaller)", configs.kolmogorov(n=16), 2),
             ("channel 64x32x16 mixed", mixed, 3),
             ("taylor-green 8^3 (smaller)", configs.taylor_green(n=8), 1)]
    for label, case, steps in cases:
        print("static diagnostics:", label, flush=True)
        g = api.load_case(case)
        g.step(steps)
        ra, a, rb, b = both(g)
        assert ra == 0 and rb == 0, (ra, rb)
        assert a == b, (label, a, b)
        assert a["step"] == steps and a["ke"] > 0.0
        if case.markers is not None:
            assert a["drag"] != 0.0
        g.close()


def test_static_diagnostics_state_errors():
    g = api.FenepGPU(configs.kolmogorov(n=16).params)
    d = FpDiag()
    assert api.lib().fp_diagnostics_static(g.h, C.byref(d)) == 3   # FP_E_STATE before set_state
    assert api.lib().fp_diagnostics_static(None, C.byref(d)) == 1
    g.close()
