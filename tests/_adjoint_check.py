"""Discrete-adjoint checks, run in a fresh interpreter (torch first, so that
the C-ABI library binds to the HIP runtime torch loaded).

  python tests/_adjoint_check.py vjp     one-step state VJP vs central FD
  python tests/_adjoint_check.py params  12-step parameter/force gradients vs FD
  python tests/_adjoint_check.py cfg5    64x64x32, 400 steps, checkpoints every
                                         20, drag averaged over the last 200
                                         steps; torch.autograd path vs FD

Finite differences replay the selections of the recorded run (the same
discrete C+- branch the adjoint differentiates).  Each check prints one JSON
line with the relative errors.
"""
import json
import math
import os
import sys
import time

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import numpy as np  # noqa: E402

from fortran8_amd import api, configs  # noqa: E402
from fortran8_amd.adjoint import DragRun, drag_objective  # noqa: E402
from fortran8_amd.fields import State, interior, interior_c, pad_periodic  # noqa: E402

DEV = torch.device("cuda", 0)


def flat_state(st):
    """(16, nz, ny, nx): U V W P C(6) A0(3) T0(3)"""
    planes = [interior(st.U), interior(st.V), interior(st.W), interior(st.P)]
    c = interior_c(st.C)
    planes += [c[..., q] for q in range(6)]
    planes += [st.hist[q] for q in range(6)]
    return np.stack(planes)


def make_state(x, b0):
    """ghosted State from (16, nz, ny, nx) and the B0 history (3, nz, ny, nx)"""
    C = np.stack([pad_periodic(x[4 + q]) for q in range(6)], axis=-1)
    hist = np.concatenate([x[10:16], b0]).copy()
    return State(pad_periodic(x[0]), pad_periodic(x[1]), pad_periodic(x[2]), pad_periodic(x[3]), C, hist, 1)


def small_case(**kw):
    case = configs.adjoint_case(nx=32, ny=32, nz=16, **kw)
    g = api.load_case(case)
    g.step(4)   # developed histories and flow
    st = g.get_state()
    return case, g, st


def check_vjp(**kw):
    case, g, st = small_case(**kw)
    x0 = flat_state(st)
    b0 = st.hist[6:9].copy()
    rng = np.random.default_rng(7)
    scale = np.array([max(np.std(p), 1e-3) for p in x0])[:, None, None, None]
    w = rng.standard_normal(x0.shape) * scale
    vbar = rng.standard_normal(x0.shape)
    wd = 0.7
    # base recording and adjoint
    g.set_state(make_state(x0, b0))
    g.adj_record(1, 1)
    vb = torch.from_numpy(vbar.reshape(-1)).to(DEV)
    bi = torch.zeros(vb.numel(), dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    g.adj_backward(weights=[wd], bar_final=vb.data_ptr(), bar_initial=bi.data_ptr())
    adj = float(np.dot(bi.cpu().numpy(), w.reshape(-1)))

    def J(x):
        g.set_state(make_state(x, b0))
        d = g.adj_record(1, 1, replay=True)
        return float(np.sum(vbar * flat_state(g.get_state()))) + wd * d[0]

    out = {"check": "vjp", "params": {k: float(v) for k, v in kw.items()}}
    for h in (1e-5, 1e-6):
        fd = (J(x0 + h * w) - J(x0 - h * w)) / (2 * h)
        out["fd_h%g" % h] = fd
        out["rel_h%g" % h] = abs(fd - adj) / abs(adj)
    out["adjoint"] = adj
    # per-block attribution: only the velocity planes / only C / only the
    # histories (which enter linearly: a large step is exact)
    for name, sl, h in (("uvwp", slice(0, 4), 1e-5), ("c", slice(4, 10), 1e-5), ("hist", slice(10, 16), 1e-2)):
        wm = np.zeros_like(w)
        wm[sl] = w[sl]
        a = float(np.dot(bi.cpu().numpy(), wm.reshape(-1)))
        fd = (J(x0 + h * wm) - J(x0 - h * wm)) / (2 * h)
        out["rel_" + name] = abs(fd - a) / max(abs(a), 1e-300)
    print(json.dumps(out), flush=True)


def check_vjp_smac_weight():
    """The one-step dot-product test with smac.f90's spreading weight
    dV = pi DC/NL dX^2 in place of dX^3 (smac.f90:1129, 1428)."""
    from fortran8_amd.fields import cylinder_markers
    d = 2 * math.pi / 32
    nl = cylinder_markers([(math.pi, math.pi / 2)], 0.5, d, d, 16).nl
    check_vjp(ibm_weight=math.pi * 0.5 / nl * d * d)


def central(J, h):
    return (J(h) - J(-h)) / (2 * h)


def richardson(J, h):
    """central difference with the h^2 term removed: (4 D(h/2) - D(h)) / 3"""
    return (4 * central(J, h / 2) - central(J, h)) / 3


def stretch(x0, trc):
    """conformation stretched along x to tr C ~ trc (L^2 then matters)"""
    x = x0.copy()
    x[4] += trc - 3.0
    return x


def check_params():
    case, g, st = small_case(newton_eps=1e-12, newton_itrmax=50)
    x0 = stretch(flat_state(st), 30.0)
    b0 = st.hist[6:9].copy()
    nsteps, every = 12, 5
    rng = np.random.default_rng(11)
    wts = rng.uniform(0.5, 1.5, nsteps)
    p = case.params
    fx = np.stack([np.asarray(a) for a in case.force])
    dF = rng.standard_normal(fx.shape) * 0.05

    def run(wi, beta, L2, force, replay):
        g.set_physics(wi, beta, math.sqrt(L2))
        g.set_state(make_state(x0, b0))
        g.set_body_force([force[0].copy(), force[1].copy(), force[2].copy()])
        d = g.adj_record(nsteps, every, replay=replay)
        return float(np.dot(wts, d))

    base = (p.wi, p.beta, p.lp ** 2)
    J0 = run(*base, fx, False)
    fb = torch.zeros(3 * fx[0].size, dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    grads = g.adj_backward(weights=wts, bar_force=fb.data_ptr())
    dirF = float(np.dot(fb.cpu().numpy(), dF.reshape(-1)))
    out = {"check": "params", "J": J0}
    names = ("wi", "beta", "L2")
    for q in range(3):
        def Jq(t):
            x = list(base)
            x[q] += t
            return run(*x, fx, True)
        out["d" + names[q]] = grads[q]
        fd6 = central(Jq, 1e-6 * base[q])
        out["rel_%s_h1e-06" % names[q]] = abs(fd6 - grads[q]) / abs(grads[q])
        fd = richardson(Jq, 1e-4 * base[q])
        out["fd_" + names[q]] = fd
        out["rel_" + names[q]] = abs(fd - grads[q]) / abs(grads[q])
    fd = richardson(lambda t: run(*base, fx + t * dF, True), 1e-4)
    out["dF_dir"] = dirF
    out["fd_F_dir"] = fd
    out["rel_F"] = abs(fd - dirF) / abs(dirF)
    print(json.dumps(out), flush=True)


def cfg5_run(newton_eps, hrels, out, tag, timing, itrmax=10000, precision=0, fd=True):
    case = configs.adjoint_case(newton_eps=newton_eps, newton_itrmax=itrmax)   # 64x64x32, Lz = pi
    case.params.precision = precision
    nsteps, every, last = 400, 20, 200
    p = case.params
    fx = np.stack([np.asarray(a) for a in case.force])
    run = DragRun(case, nsteps, every, last)
    wi = torch.tensor(p.wi, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(p.beta, dtype=torch.float64, requires_grad=True)
    L2 = torch.tensor(p.lp ** 2, dtype=torch.float64, requires_grad=True)
    F = torch.from_numpy(fx).to(DEV).requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = drag_objective(run, wi, beta, L2, F)
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if timing:
        # plain forward of the same steps (no recording) for the cost ratios
        g = run.g
        g.set_physics(p.wi, p.beta, p.lp)
        g.set_state(case.state)
        g.set_body_force([fx[0].copy(), fx[1].copy(), fx[2].copy()])
        t3 = time.perf_counter()
        g.step(nsteps)
        t4 = time.perf_counter()
        out.update({"%s_J" % tag: float(J), "%s_t_record_s" % tag: t1 - t0, "%s_t_backward_s" % tag: t2 - t1,
                    "%s_t_forward_s" % tag: t4 - t3,
                    "%s_ratio_backward_forward" % tag: (t2 - t1) / (t4 - t3),
                    "%s_ratio_record_forward" % tag: (t1 - t0) / (t4 - t3),
                    "%s_ratio_record_backward_forward" % tag: (t2 - t0) / (t4 - t3)})
    grads = [float(wi.grad), float(beta.grad), float(L2.grad)]
    rng = np.random.default_rng(5)
    dF = rng.standard_normal(fx.shape)
    dirF = float(torch.sum(F.grad.cpu() * torch.from_numpy(dF)))
    out["%s_dF_dir" % tag] = dirF
    for q, name in enumerate(("wi", "beta", "L2")):
        out["%s_d%s" % (tag, name)] = grads[q]
    if not fd:
        run.close()
        return
    run.replay = True
    base = [p.wi, p.beta, p.lp ** 2]
    for q, name in enumerate(("wi", "beta", "L2")):
        def Jq(t):
            x = list(base)
            x[q] += t
            return run.forward(*x, fx)[0]
        for hrel in hrels:
            fd = central(Jq, hrel * base[q])
            out["%s_fd_%s_h%g" % (tag, name, hrel)] = fd
            out["%s_rel_%s_h%g" % (tag, name, hrel)] = abs(fd - grads[q]) / abs(grads[q])
        fd = richardson(Jq, 1e-4 * base[q])
        out["%s_fd_%s_rich" % (tag, name)] = fd
        out["%s_rel_%s_rich" % (tag, name)] = abs(fd - grads[q]) / abs(grads[q])
    fd = central(lambda t: run.forward(*base, fx + t * dF)[0], 1e-6)
    out["%s_rel_F" % tag] = abs(fd - dirF) / abs(dirF)
    run.close()


def check_cfg5():
    out = {"check": "cfg5"}
    # the reference's Newton tolerance (1e-6): the stopping rule makes the
    # step piecewise smooth in the parameters
    cfg5_run(1e-6, (1e-4, 1e-6), out, "ref", True)
    # Newton run to a fixed count of 8 iterations (tolerance 0): converged to
    # round-off and smooth in the parameters; central differences at h = 1e-6
    cfg5_run(0.0, (1e-6,), out, "conv", False, itrmax=8)
    # mixed-precision forward (FP32 bulk), FP64 reverse sweep along the
    # recomputed segments: its gradients against the FP64 ones
    cfg5_run(1e-6, (), out, "mixed", True, precision=1, fd=False)
    for name in ("wi", "beta", "L2"):
        out["mixed_rel_%s_vs_fp64" % name] = abs(out["mixed_d" + name] - out["ref_d" + name]) / abs(out["ref_d" + name])
    out["mixed_rel_F_vs_fp64"] = abs(out["mixed_dF_dir"] - out["ref_dF_dir"]) / abs(out["ref_dF_dir"])
    # complex-step gradients of the FP64 CPU oracle (tests/golden/
    # make_complex_step.py): the same objective, Newton settings and force
    # direction
    with open(os.path.join(ROOT, "tests", "golden", "cfg5_complex_step.json")) as f:
        cs = json.load(f)
    for tag in ("conv", "ref"):
        for name in ("wi", "beta", "L2"):
            a, c = out["%s_d%s" % (tag, name)], cs[tag]["d" + name]
            out["%s_rel_%s_cs" % (tag, name)] = abs(a - c) / abs(c)
        a, c = out["%s_dF_dir" % tag], cs[tag]["dF"]
        out["%s_rel_F_cs" % tag] = abs(a - c) / abs(c)
        out["%s_J_cs" % tag] = cs[tag]["J"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "vjp"
    print("adjoint check:", which, flush=True)
    {"vjp": check_vjp, "vjp_smac_weight": check_vjp_smac_weight, "params": check_params,
     "cfg5": check_cfg5}[which]()
