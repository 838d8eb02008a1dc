"""Complex-step gradients of the sensitivity configuration's objective, from
the complex-step build of the FP64 CPU oracle (oracle/liboracle_cs.so), for
tests/test_gpu_adjoint.py.

Configuration: 64x64x32 (Lz = pi), Kolmogorov forcing, one cylinder, 400
steps; J = drag averaged over the last 200 steps.  Directions: Wi, beta, L^2
and the body-force field along dF = standard normal draws of
numpy.random.default_rng(5) (shape (3, NZ, NY, NX)).  Two Newton settings:
"conv" (tolerance 0, 8 iterations: converged to round-off) and "ref" (the
reference's tolerance 1e-6).

    python tests/golden/make_complex_step.py   # writes cfg5_complex_step.json

Runs the eight evaluations in parallel, one thread each (about 15 minutes on
8 cores)."""
import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
H = 1e-30
NSTEPS, LAST = 400, 200


def run(job):
    os.environ["OMP_NUM_THREADS"] = "1"
    sys.path.insert(0, ROOT)
    import numpy as np
    from fortran8_amd import configs
    from oracle import pyoracle as po
    tag, direction = job
    eps, itrmax = (0.0, 8) if tag == "conv" else (1e-6, 10000)
    case = configs.adjoint_case(newton_eps=eps, newton_itrmax=itrmax)
    cs = po.ComplexStep(case)
    if direction == "F":
        fx = np.stack([np.asarray(a) for a in case.force])
        dF = np.random.default_rng(5).standard_normal(fx.shape)
        cs.seed_force(dF, H)
    else:
        cs.seed_param({"wi": 0, "beta": 1, "L2": 2}[direction], H)
    re, im = cs.drag_series(NSTEPS)
    w = np.zeros(NSTEPS)
    w[NSTEPS - LAST:] = 1.0 / LAST
    return tag, direction, float(np.dot(w, re)), float(np.dot(w, im)) / H


def main():
    jobs = [(t, d) for t in ("conv", "ref") for d in ("wi", "beta", "L2", "F")]
    with Pool(len(jobs)) as pool:
        res = pool.map(run, jobs)
    out = {"config": "adjoint_case 64x64x32, 400 steps, J = mean drag of the last 200",
           "complex_step": H, "force_direction": "numpy.random.default_rng(5).standard_normal((3, NZ, NY, NX))"}
    for tag, d, J, g in res:
        out.setdefault(tag, {"J": J})["d" + d] = g
        assert out[tag]["J"] == J
    path = os.path.join(HERE, "cfg5_complex_step.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
