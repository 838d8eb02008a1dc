"""C-ABI library checks that need no GPU: it loads, exports every entry
point include/fenep_gpu.h declares, validates parameters before touching
the device, and produces the seeded input stream the host code uses."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from fortran8_amd import api, fields
from fortran8_amd.params import FpParams, default_params


def test_library_exports_every_declared_entry_point():
    names = api.header_exports()
    assert {"fp_create", "fp_step", "fp_last_error", "fp_seeded_uniform"} <= set(names) and len(names) >= 17
    L = api.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert (" T " + n) in out, n


def test_library_is_built_for_gfx950():
    out = subprocess.run(["/opt/rocm/lib/llvm/bin/llvm-objdump", "--offloading", api.LIB_PATH],
                         capture_output=True, text=True)
    assert "gfx950" in (out.stdout + out.stderr)


def test_default_params_match_reference_constants():
    p = FpParams()
    assert api.lib().fp_default_params(C.byref(p)) == 0
    assert p.newton_eps == 1e-6 and p.newton_itrmax == 10000 and p.newton_dc == 1e-5   # fenep.f90:526-528
    assert p.sor_eps == 1e-5 and p.sor_itrmax == 10000                                  # fenep.f90:856-857
    q = default_params()
    for name in ("nx", "ny", "nz", "newton_eps", "newton_dc", "sor_eps", "method"):
        assert getattr(p, name) == getattr(q, name)


@pytest.mark.parametrize("bad", [dict(nx=12), dict(ny=2), dict(dt=-1.0), dict(re=0.0), dict(method=5),
                                 dict(precision=3), dict(nz=2048)])
def test_create_rejects_invalid_parameters_without_a_device(bad):
    p = default_params(**bad)
    h = C.c_void_p()
    assert api.lib().fp_create(C.byref(p), C.byref(h)) == 1
    assert not h.value


def test_null_handles_are_safe():
    L = api.lib()
    assert L.fp_destroy(None) == 0
    assert L.fp_last_error(None) == b"null context"
    assert L.fp_step(None, 1, 1) == 1
    buf = (C.c_ubyte * 8)()
    assert L.fp_get_selection(None, buf) == 1
    out3 = (C.c_double * 3)()
    assert L.fp_check_selection(None, out3) == 1


def test_seeded_uniform_stream_matches_host_implementation():
    a = api.seeded_uniform(12345, 1000, offset=77)
    b = fields.seeded_uniform(12345, 1000, offset=77)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.05


def test_oracle_is_not_linked_by_the_product():
    out = subprocess.run(["ldd", api.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", api.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_library_matches_the_sources_of_this_tree():
    from fortran8_amd import build
    assert api.build_id() == build.source_hash()
    assert len(api.build_id()) == 16


def test_params_and_diag_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of fp_params and fp_diag have the header's field
    offsets and sizes (gcc on the plain-C header)."""
    from fortran8_amd.params import FpDiag
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fenep_gpu.h"', "int main(void) {"]
    for st, cls in (("fp_params", FpParams), ("fp_diag", FpDiag)):
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (st, st))
        for name, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (st, name, st, name))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", header, str(src), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for st, cls in (("fp_params", FpParams), ("fp_diag", FpDiag)):
        assert got[(st, "size")] == C.sizeof(cls), st
        for name, _ in cls._fields_:
            assert got[(st, name)] == getattr(cls, name).offset, (st, name)
