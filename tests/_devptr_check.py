"""Device-pointer boundary round trip, run in a fresh interpreter: torch is
imported and initialised first so that the C-ABI library binds to the HIP
runtime torch already loaded (both use the soname libamdhip64.so.7)."""
import os
import sys

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import numpy as np  # noqa: E402

from fortran8_amd import api, configs  # noqa: E402
from fortran8_amd.fields import interior, interior_c  # noqa: E402

case = configs.kolmogorov(n=16)
g = api.load_case(case)
g.step(2)
n = 16 ** 3
dev = torch.device("cuda", 0)
bufs = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(4)] + \
       [torch.zeros(6 * n, dtype=torch.float64, device=dev)]
assert api.lib().fp_get_state_device(g.h, *[b.data_ptr() for b in bufs], None) == 0
torch.cuda.synchronize()
st = g.get_state()
assert np.array_equal(bufs[0].cpu().numpy().reshape(16, 16, 16), interior(st.U))
g2 = api.FenepGPU(case.params)
assert api.lib().fp_set_state_device(g2.h, *[b.data_ptr() for b in bufs], 0, None) == 0
st2 = g2.get_state()
assert np.array_equal(interior(st2.V), interior(st.V))
assert np.array_equal(interior_c(st2.C), interior_c(st.C))
print("device pointer round trip ok")
