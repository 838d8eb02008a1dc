"""Throughput benchmark of the FENE-P + IBM step on one MI355X.

Workload (N=1): 256x128x128 cells, FENE-P + 8 IBM cylinders, y-walls, FFT
Poisson (the throughput configuration).  A "step" is one full time step of
the state resident in HBM.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--precision mixed|f64]
                  [--no-cpu] [--cpu-steps S] [--no-f64] [--no-adjoint] [--small]

The roofline kernel's duration comes from HIP events recorded around it on
its stream inside the timed region (the only events there: every phase
timed would add ~7 % to the step); the per-phase breakdown comes from a few
steps run afterwards with every phase timed.  HBM traffic per launch of the
roofline kernel is the rocprofv3 FETCH_SIZE/WRITE_SIZE reading kept in
profiles/traffic.txt (already scaled by the copy-kernel calibration), and the
VALU instruction count of the selection kernel its SQ_INSTS_VALU reading.

Beside the mixed-precision step the same line carries the FP64 step timed
over the same steps (the comparator of the mixed/FP64 speed), the accuracy of
the mixed state against the FP64 one at the end of those runs, and the cost
of the discrete adjoint on the 64x64x32 sensitivity configuration.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from fortran8_amd import configs, api  # noqa: E402

# Algorithmic HBM words per cell-step of the canonical pass list (FP64 plain:
# 80 words = 640 B).
STEP_WORDS = 80
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of the single-kernel phases: (words read +
# written at the storage width, extra bytes), per precision (0 FP64, 1
# mixed).  select: C 6 read, the selection byte written.  cstar_newton: C 6 +
# U,V,W read, selection byte read, C 6 + Peterlin f written, Newton count
# (int32) written.  predictor: U,V,W,P, C 6, f, A0 3, T0 3 read; A0 3, T0 3
# written, and the IBM predictor's outputs: FP64 u~ 3 and the ibm_predict
# right-hand side 3, mixed the increment d_q 3 alone.
KERNEL_WORDS = {
    ("fenep_select", 0): (6, 1), ("fenep_select", 1): (6, 1),
    ("fenep_cstar_newton", 0): (16, 5), ("fenep_cstar_newton", 1): (16, 5),
    ("predictor", 0): (29, 0), ("predictor", 1): (26, 0),
}


def cells_per_launch(p, prec, phase, phases):
    """Cells one launch of the phase's kernel processes.  Each timed phase is
    one kernel.  With y-walls the per-cell mixed predictor, which runs where
    the z-streamed predictor's tiling does not fit the grid, leaves the 2
    wall rows to its own instance (timed as predictor_walls); every other
    kernel covers all cells."""
    cells = p.nx * p.ny * p.nz
    if prec == 1 and p.walls_y and phase == "predictor" and "predictor_walls" in phases:
        return p.nx * (p.ny - 2) * p.nz
    return cells


def mixed_step_bytes(n_dc):
    """Algorithmic HBM bytes per cell of the mixed IBM step with y-walls at
    the throughput size (4 B words, FP64 where stored so): select 25 (C read,
    selection byte written), CStar/Newton 69, predictor 104 (its output is
    the increment d_q alone), the two Helmholtz sweeps in one pass 24 (d_q
    read, d written), correction 52 (u^n, d, P, Phi read, u, P written);
    Poisson: divergence fused into the forward x transform 28 (u^n and d
    read, half spectrum written), the kx-slab solve 8 (half spectrum read and
    written), inverse x transform with the FP64 accumulation of Phi 12; the
    FP64 defect check every 8 steps 3.5 on average (b written, residual 20).
    Each further defect-correction sweep adds 36.  The sparse IBM passes and
    the FP64 re-decision of selections are not counted."""
    return 322.0 + 3.5 + 36.0 * max(0, n_dc - 1)


def run_gpu(case, steps, warmup, keep_state=False):
    """Timed region: `steps` steps with HIP events around the predictor (the
    roofline kernel) of every 8th step (timing level 3: each event idles the
    stream for a few microseconds); its average launch duration is taken over
    those sampled launches.  The state after the timed region is kept (for
    the accuracy figures) before, outside the timed region, a few steps run
    with every phase timed for the breakdown."""
    g = api.load_case(case)
    g.enable_timing(3)
    g.step(warmup, sync=True)           # warm-up (also fills the event pool)
    g.enable_timing(3)
    s0 = g.diagnostics()["step"]
    t0 = time.perf_counter()
    g.step(steps, sync=True)
    t1 = time.perf_counter()
    sampled = max(1, sum(1 for s in range(s0, s0 + steps) if s % 8 == 0))
    kt_timed = {k: v / sampled for k, v in g.kernel_times().items()}
    d = g.diagnostics()
    st = g.get_state() if keep_state else None
    nb = max(1, min(steps, 10))
    g.enable_timing(1)
    g.step(nb, sync=True)
    kt = {k: v / nb for k, v in g.kernel_times().items()}
    g.enable_timing(0)
    g.close()
    wall = (t1 - t0) / steps
    return wall, kt_timed, kt, d, st


def accuracy(ref, dref, got, dgot):
    """Relative L2 of u (stacked), U, V, W, P and C, and the relative drag and
    kinetic-energy differences, of `got` against `ref`."""
    from fortran8_amd.fields import interior, interior_c, rel_l2
    e = {n: rel_l2(interior(getattr(got, n)), interior(getattr(ref, n))) for n in ("U", "V", "W", "P")}
    e["C"] = rel_l2(interior_c(got.C), interior_c(ref.C))
    e["u"] = rel_l2(np.concatenate([interior(getattr(got, n)).ravel() for n in "UVW"]),
                    np.concatenate([interior(getattr(ref, n)).ravel() for n in "UVW"]))
    e["drag"] = abs(dgot["drag"] - dref["drag"]) / abs(dref["drag"])
    e["ke"] = abs(dgot["ke"] - dref["ke"]) / abs(dref["ke"])
    return {k: float("%.3e" % v) for k, v in e.items()}


def adjoint_cost(reps=2):
    """Cost of the discrete adjoint on the sensitivity configuration (64x64x32,
    one cylinder, 400 FP64 steps, checkpoints every 20, drag averaged over the
    last 200): wall time of a plain forward run, of the recording run, and of
    the reverse sweep (which recomputes each checkpoint segment), each the
    shortest of `reps` runs."""
    case = configs.adjoint_case()
    nsteps, every, last = 400, 20, 200
    g = api.load_case(case)
    w = np.zeros(nsteps)
    w[nsteps - last:] = 1.0 / last
    g.step(5, sync=True)                 # warm-up
    fwd = rec = rev = float("inf")
    for _ in range(reps):
        g.set_state(case.state)
        t0 = time.perf_counter()
        g.step(nsteps, sync=True)
        t1 = time.perf_counter()
        g.set_state(case.state)
        t2 = time.perf_counter()
        g.adj_record(nsteps, every)
        t3 = time.perf_counter()
        grads = g.adj_backward(weights=w)
        t4 = time.perf_counter()
        fwd, rec, rev = min(fwd, t1 - t0), min(rec, t3 - t2), min(rev, t4 - t3)
    g.close()
    return {"config": "64x64x32, 400 steps, checkpoints every 20", "runs": reps, "forward_s": round(fwd, 4),
            "record_s": round(rec, 4), "reverse_s": round(rev, 4),
            "reverse_over_forward": round(rev / fwd, 3),
            "record_and_reverse_over_forward": round((rec + rev) / fwd, 3),
            "dJ_dWi": grads[0], "dJ_dbeta": grads[1]}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def read_profile_counts(kind, precision):
    """Per-launch readings of profiles/traffic.txt: lines `traffic <phase>
    <precision> <bytes>` (calibrated HBM bytes) or `valu <phase> <precision>
    <instructions>` (SQ_INSTS_VALU, wave instructions per dispatch)."""
    out = {}
    path = os.path.join(HERE, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return out
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == kind and f[2] == precision:
            out[f[1]] = float(f[3])
    return out


# VALU issue: a wave64 VALU instruction occupies its SIMD for 4 cycles (16
# lanes per cycle, non-packed FP32/FP64); MI355X has 256 CUs x 4 SIMDs and a
# 2.4 GHz peak engine clock.
SIMDS, CLOCK_HZ, CYCLES_PER_VALU = 1024, 2.4e9, 4


def valu_issue_ms(valu_insts):
    return valu_insts * CYCLES_PER_VALU / SIMDS / CLOCK_HZ * 1e3


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
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=["mixed", "f64"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--small", action="store_true", help="64x32x32 smoke size")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 comparator and the accuracy figures")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost figures")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the whole domain runs on a single MI355X "
              "(the library has no inter-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    prec = 1 if a.precision == "mixed" else 0
    if a.small:
        case = configs.throughput(nx=64, ny=32, nz=32, precision=prec, ncyl=1)
    else:
        case = configs.throughput(precision=prec)
    compare = prec == 1 and not a.no_f64
    wall, kt_timed, kt, d, st = run_gpu(case, a.steps, a.warmup, keep_state=compare)
    cells = case.cells
    value = cells / wall / 1e6
    width = 8 if prec == 0 else 4
    # The HBM roofline is reported for the predictor, the largest kernel that
    # is bound by memory: the selection and CStar/Newton kernels, of similar
    # duration, are bound by VALU issue (listed under compute_bound; in the
    # mixed step the selection also runs beside the Poisson solve on a second
    # stream, so its duration there includes the sharing).
    dom = "predictor"
    words, extra = KERNEL_WORDS[(dom, prec)]
    units = cells_per_launch(case.params, prec, dom, kt)
    dom_bytes = (words * width + extra) * units
    dom_ms = kt_timed.get(dom, 0.0)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else None
    traffic = read_profile_counts("traffic", a.precision).get(dom)
    step_bytes = STEP_WORDS * 8 * cells if prec == 0 else mixed_step_bytes(d.get("n_dc", 0)) * cells
    walls = "true" if case.params.walls_y else "false"
    mixed_pred = "fp::m_predictor<false>" if "predictor_walls" in kt else "fp::m_predictor_z<%s>" % walls
    # the selection streamed along z where a block holds whole x-rows
    streamed = case.params.nx in (64, 128, 256) and case.params.nz % 4 == 0
    kernel_names = {"fenep_select": ("fp::m_select_z" if streamed else "fp::m_select") if prec else "fp::k_select",
                    "fenep_cstar_newton": "fp::m_cstar_newton" if prec else "fp::k_cstar_newton<%s>" % walls,
                    "predictor": mixed_pred if prec else "fp::k_predictor<%s>" % walls}
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
        # phase breakdown from steps run after the timed region with every
        # phase timed (the timed region times the predictor only)
        "phases_ms": {k: round(v, 4) for k, v in kt.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dom],
            "algorithmic_bytes_per_cell": words * width + extra,
            "cells_per_launch": units,
            "launch_ms": round(dom_ms, 4),
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            # the kernel with the largest share of the step: bound by VALU
            # issue, not by HBM; its fraction is the VALU issue time of its
            # instructions (SQ_INSTS_VALU) over its duration
            "dominant": dominant_valu(kt, kernel_names, a.precision),
        },
        "step_roofline": {
            "algorithmic_bytes_per_cell_step": step_bytes / cells,
            "achieved": round(step_bytes / wall / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(step_bytes / wall / 1e9 / HBM_PEAK_GBS, 4),
        },
        "diag": {k: d[k] for k in ("ke", "cfl", "trc_min", "trc_max", "drag", "div_max", "newton_itr_mean", "n_dc",
                                    "dc_residual", "sel_guarded", "nan_flag")},
    }
    if compare:
        # the FP64 step over the same steps from the same state: its speed is
        # the comparator of the mixed step, its state the accuracy reference
        ref = configs.throughput(precision=0)
        wall64, _, kt64, d64, st64 = run_gpu(ref, a.steps, a.warmup, keep_state=True)
        out["fp64_comparator"] = {"ms_per_step": round(wall64 * 1e3, 4),
                                  "value": round(cells / wall64 / 1e6, 3), "unit": "Mcell-steps/s",
                                  "mixed_speedup": round(wall64 / wall, 3)}
        # floor: the FP64 step from the FP32-rounded initial state
        fl = configs.throughput(precision=0)
        for nm in ("U", "V", "W", "P", "C"):
            setattr(fl.state, nm, getattr(fl.state, nm).astype(np.float32).astype(np.float64))
        gf = api.load_case(fl)
        gf.step(a.warmup + a.steps, sync=True)
        stf, dfl = gf.get_state(), gf.diagnostics()
        gf.close()
        out["accuracy"] = {"steps": a.warmup + a.steps, "reference": "FP64 GPU step (bit-identical to the oracle)",
                           "mixed": accuracy(st64, d64, st, d),
                           "floor_fp32_initial_state": accuracy(st64, d64, stf, dfl)}
    if not a.no_adjoint and not a.small:
        out["adjoint"] = adjoint_cost()
    if not a.no_cpu:
        cv, cores, sec = cpu_baseline(configs.throughput(precision=0), a.cpu_steps)
        out["cpu_baseline"] = {"value": round(cv, 4), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
                               "cpu": cpu_model(),
                               "sample": "%d step(s) of the full 256x128x128 case on the FP64 CPU restatement "
                                         "(OpenMP, %d threads), %.1f s" % (a.cpu_steps, cores, sec)}
    print(json.dumps(out), flush=True)


def dominant_valu(kt, kernel_names, precision):
    """The selection kernel (the largest share of the mixed step's kernel
    time) against its VALU issue time."""
    phase = "fenep_select"
    insts = read_profile_counts("valu", precision).get(phase)
    ms = kt.get(phase)
    if not insts or not ms:
        return None
    floor = valu_issue_ms(insts)
    return {"kernel": kernel_names[phase], "bound": "valu", "valu_insts": insts,
            "issue_ms": round(floor, 4), "launch_ms": round(ms, 4), "frac": round(floor / ms, 4),
            "note": "launch_ms from the phase-timed steps; the kernel shares the GPU with the Poisson solve"}


if __name__ == "__main__":
    main()
