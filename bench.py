"""Throughput benchmark of the FENE-P + IBM step on one MI355X.

Workload (N=1): 256x128x128 cells, FENE-P + 8 IBM cylinders, y-walls, FFT
Poisson (the throughput configuration).  A "step" is one full time step of
the state resident in HBM.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--precision mixed|f64]
                  [--no-cpu] [--small]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from fortran8_amd import configs, api  # noqa: E402

# Algorithmic HBM words per cell-step of the canonical pass list (FP64 plain:
# 80 words = 640 B).  Pass P1 (FENE-P: C6 + UVW read, C6 write) is 15 words.
STEP_WORDS = 80
FENEP_WORDS = 15
HBM_PEAK_GBS = 8000.0


def mixed_step_bytes(n_dc):
    return 320.0 + 60.0 * n_dc


def run_gpu(case, steps, warmup):
    g = api.load_case(case)
    g.step(warmup, sync=True)
    t0 = time.perf_counter()
    g.step(steps, sync=True)
    t1 = time.perf_counter()
    wall = (t1 - t0) / steps
    # phase timing with HIP events on the context's stream (separate pass)
    g.enable_timing(True)
    g.step(steps, sync=True)
    kt = {k: v / steps for k, v in g.kernel_times().items()}
    g.enable_timing(False)
    d = g.diagnostics()
    g.close()
    return wall, kt, d


def cpu_baseline(case, steps=1):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    from oracle import pyoracle as po
    o = po.load_case(case)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return case.cells * steps / dt / 1e6, cores, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["mixed", "f64"], default="f64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="64x32x32 smoke size")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the whole domain runs on a single MI355X "
              "(no inter-GPU path in this build); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    prec = 1 if a.precision == "mixed" else 0
    if a.small:
        case = configs.throughput(nx=64, ny=32, nz=32, precision=prec, ncyl=1)
    else:
        case = configs.throughput(precision=prec)
    wall, kt, d = run_gpu(case, a.steps, a.warmup)
    cells = case.cells
    value = cells / wall / 1e6
    fen_ms = kt.get("fenep_cstar_newton", 0.0) + kt.get("fenep_select", 0.0)
    fen_bytes = FENEP_WORDS * (8 if prec == 0 else 4) * cells
    step_bytes = STEP_WORDS * 8 * cells if prec == 0 else mixed_step_bytes(d.get("n_dc", 0)) * cells
    out = {
        "metric": "Mcell-steps/s 256x128x128 FENE-P+IBM",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32+f64",
        "data": "synthetic (seeded noise ICs, seeded cylinder layout)",
        "config": {"workload": case.name, "grid": [case.params.nx, case.params.ny, case.params.nz],
                   "cylinders": len(case.markers.centers), "precision": a.precision,
                   "poisson": "fft+tridiagonal", "parallelism": "single GPU"},
        "phases_ms": {k: round(v, 4) for k, v in kt.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "fenep (select + cstar/newton)",
            "achieved": round(fen_bytes / (fen_ms * 1e-3) / 1e9, 2) if fen_ms > 0 else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(fen_bytes / (fen_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if fen_ms > 0 else None,
            "traffic": None,
        },
        "step_roofline": {
            "algorithmic_bytes_per_cell_step": step_bytes / cells,
            "achieved": round(step_bytes / wall / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(step_bytes / wall / 1e9 / HBM_PEAK_GBS, 4),
        },
        "diag": {k: d[k] for k in ("ke", "cfl", "trc_min", "trc_max", "drag", "div_max", "newton_itr_mean", "n_dc",
                                    "dc_residual", "sel_guarded", "nan_flag")},
    }
    if not a.no_cpu:
        cv, cores, sec = cpu_baseline(configs.throughput(precision=0), a.cpu_steps)
        out["cpu_baseline"] = {"value": round(cv, 4), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
                               "sample": "%d step(s) of the full 256x128x128 case on the FP64 CPU restatement "
                                         "(OpenMP), %.1f s" % (a.cpu_steps, sec)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
