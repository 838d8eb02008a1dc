"""The oracle's pins must fail plausible restatement mistakes (CPU only).

Each case copies oracle/fenep_oracle.c to a temporary directory, plants one
mistake there, builds that copy with the oracle's own flags, and runs the pin
tests against it (FENEP_ORACLE_LIB).  The case passes when the pin tests fail.
The oracle in the tree is never modified."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "fenep_oracle.c")

# (name, original text, mistaken text, pin tests expected to catch it)
MUTATIONS = [
    ("cstar_x_faces_swapped",      # fenep.f90:499 (Cp and Cn of the x faces exchanged)
     "s3 * o->Cp[0][q][xl] + s1 * o->Cp[0][q][xr]) + s4 * o->Cn[0][q][xl]) + s2 * o->Cn[0][q][xr])",
     "s3 * o->Cn[0][q][xl] + s1 * o->Cn[0][q][xr]) + s4 * o->Cp[0][q][xl]) + s2 * o->Cp[0][q][xr])",
     "upwind or without_flow"),
    ("fft_predictor_T0_by_T",      # fenep.f90:1473
     "+ (dt * o->B[c][m]) / 2.0)\n               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)",
     "+ (dt * o->B[c][m]) / 2.0)\n               + (dt * (o->T[c][m] + o->T[c][m])) / 2.0)",
     "history"),
    ("ibm_predictor_T0_by_T",      # fenep.f90:2000
     "+ dt * o->B[c][m])\n               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)",
     "+ dt * o->B[c][m])\n               + (dt * (o->T[c][m] + o->T[c][m])) / 2.0)",
     "history"),
    ("explicit_predictor_T0_by_T",  # fenep.f90:805
     "(o->A0[c][m] + o->B0[c][m]))) / 2.0)\n               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)",
     "(o->A0[c][m] + o->B0[c][m]))) / 2.0)\n               + (dt * (o->T[c][m] + o->T[c][m])) / 2.0)",
     "history"),
    ("cn_pressure_update_without_laplacian",   # fenep.f90:1592-1595
     "o->P[g] = (o->P[g] + Ph[g]) - coef * lap;",
     "o->P[g] = (o->P[g] + Ph[g]) - 0.0 * coef * lap;",
     "balanced"),
    ("explicit_pressure_update_halved",        # fenep.f90:937
     "o->P[g] = o->P[g] + Ph[g];",
     "o->P[g] = o->P[g] + 0.5 * Ph[g];",
     "balanced"),
    ("pressure_gradient_sign",     # fenep.f90:1469 (dt*(-P(i)+P(i+1))/dX)
     "return (o->p.dt * (-p0 + p1)) / h;",
     "return (o->p.dt * (p0 - p1)) / h;",
     "balanced"),
]


@pytest.mark.parametrize("name,orig,mutant,selector", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pin_tests_catch_the_mistake(name, orig, mutant, selector, tmp_path):
    text = open(SRC).read()
    assert text.count(orig) == 1, "the original text must occur exactly once"
    csrc = tmp_path / "fenep_oracle.c"
    csrc.write_text(text.replace(orig, mutant))
    so = tmp_path / "liboracle.so"
    inc = ["-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "include")]
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-std=c11",
                           "-w", "-shared", "-o", str(so), str(csrc), "-lm"] + inc)
    env = dict(os.environ, FENEP_ORACLE_LIB=str(so), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-q",
                        "-x", "-p", "no:cacheprovider", "-k", selector], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    tail = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr[-500:]
    print(name, "->", tail)
    assert r.returncode == 1 and "failed" in tail, (name, tail)
