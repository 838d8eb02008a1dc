"""Accuracy of the mixed-precision step against the FP64 step, and the two
floors it is compared with, run in a fresh interpreter (torch is initialised
first, so the C-ABI library binds to the HIP runtime torch loaded):

- floor_ic: the FP64 step started from the FP32-rounded initial state;
- floor_storage: the FP64 step whose U, V, W, P, C are rounded to FP32 after
  every step (every operation in FP64, FP32 storage only).

Usage: _storage_floor.py <config> <steps> [out.json]; config is one of
throughput, channel, channel_sor, kolmogorov64.  Prints one JSON object."""
import json
import os
import sys

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import numpy as np  # noqa: E402

from fortran8_amd import api, configs  # noqa: E402
from fortran8_amd.fields import interior, interior_c, rel_l2  # noqa: E402
from fortran8_amd.params import POISSON_SOR, PREC_MIXED  # noqa: E402

BUILD = {"throughput": lambda: configs.throughput(),
         "channel": lambda: configs.channel_cylinder(),
         "channel_sor": lambda: configs.channel_cylinder(poisson_variant=POISSON_SOR),
         "kolmogorov64": lambda: configs.kolmogorov(n=64)}


def errors(ref, dref, got, dgot):
    """Relative L2 of each field and of the stacked velocity; V, W and the
    pressure also against the velocity scale (|dV| / |u|, and the pressure
    against u_max^2), which stays meaningful where a component is near zero;
    drag and kinetic energy relative."""
    e = {n: rel_l2(interior(getattr(got, n)), interior(getattr(ref, n))) for n in ("U", "V", "W", "P")}
    e["C"] = rel_l2(interior_c(got.C), interior_c(ref.C))
    vel_r = np.concatenate([interior(getattr(ref, n)).ravel() for n in "UVW"])
    vel_g = np.concatenate([interior(getattr(got, n)).ravel() for n in "UVW"])
    e["u"] = rel_l2(vel_g, vel_r)
    un = np.linalg.norm(vel_r)
    for n in "VW":
        e[n + "_u"] = np.linalg.norm(interior(getattr(got, n)) - interior(getattr(ref, n))) / un
    pr = interior(ref.P)
    e["P_dyn"] = np.linalg.norm(interior(got.P) - pr) / (np.sqrt(pr.size) * np.max(np.abs(vel_r)) ** 2)
    if dref.get("drag"):
        e["drag"] = abs(dgot["drag"] - dref["drag"]) / abs(dref["drag"])
    e["ke"] = abs(dgot["ke"] - dref["ke"]) / dref["ke"]
    return {k: float(v) for k, v in e.items()}


def main():
    name, steps = sys.argv[1], int(sys.argv[2])
    out = {"config": name, "steps": steps}
    print("fp64", name, steps, flush=True)
    h = api.load_case(BUILD[name]())
    h.step(steps)
    R, dR = h.get_state(), h.diagnostics()
    h.close()

    print("mixed", flush=True)
    mc = BUILD[name]()
    mc.params.precision = PREC_MIXED
    g = api.load_case(mc)
    g.step(steps)
    sel = g.check_selection()
    out["mixed"] = errors(R, dR, g.get_state(), g.diagnostics())
    out["mixed"]["sel_differ"] = sel["differ"]
    out["mixed"]["sel_fp64_fraction"] = sel["fp64"] / sel["total"]
    g.close()

    print("floor_ic", flush=True)
    c1 = BUILD[name]()
    for nm in ("U", "V", "W", "P", "C"):
        setattr(c1.state, nm, getattr(c1.state, nm).astype(np.float32).astype(np.float64))
    g = api.load_case(c1)
    g.step(steps)
    out["floor_ic"] = errors(R, dR, g.get_state(), g.diagnostics())
    g.close()

    print("floor_storage", flush=True)
    g = api.load_case(c1)
    n = c1.cells
    dev = torch.device("cuda", 0)
    bufs = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(4)] + \
           [torch.zeros(6 * n, dtype=torch.float64, device=dev)]
    L = api.lib()
    for s in range(steps):
        g.step(1)
        assert L.fp_get_state_device(g.h, *[b.data_ptr() for b in bufs], None) == 0
        for b in bufs:
            b.copy_(b.float().double())
        torch.cuda.synchronize()
        # 2: continue, keeping the potential (the SOR sweeps start from it)
        assert L.fp_set_state_device(g.h, *[b.data_ptr() for b in bufs], 2, None) == 0
        if (s + 1) % 250 == 0:
            print("  floor_storage step", s + 1, flush=True)
    out["floor_storage"] = errors(R, dR, g.get_state(), g.diagnostics())
    g.close()
    txt = json.dumps(out)
    if len(sys.argv) > 3:
        with open(sys.argv[3], "w") as f:
            f.write(json.dumps(out, indent=1))
    print(txt, flush=True)


if __name__ == "__main__":
    main()
