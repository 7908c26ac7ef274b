"""The allocation-free diagnostics entry point (sphk_diagnostics_static,
include/sphk_step.h) against sphk_diagnostics through the C-ABI: bitwise the
same energy, momentum, minimum particle timestep and mass, on two calls in a
row, the second with a smaller input (the shared per-GPU scratch keeps the
first call's partials, which the second must overwrite)."""
import ctypes as C

import numpy as np
import pytest
import torch

from research_codes_amd import _lib as L
from research_codes_amd.device import Context
from research_codes_amd.step import phys

pytestmark = pytest.mark.gpu


def _state(n, seed):
    rng = np.random.default_rng(seed)
    s = {k: rng.uniform(0.5, 1.5, n) for k in L.STATE_FIELDS}
    for k in ("vx", "vy", "vz", "divv", "g1x", "g2y", "pot"):
        s[k] = rng.normal(size=n)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in s.items()}
    return t, L.State(**{k: t[k].data_ptr() for k in L.STATE_FIELDS})


def test_static_diagnostics_equal_allocating_ones():
    lib = L.lib()
    ph = phys(c_dt=0.3, dt_max=2.0)
    with Context(0) as c:
        st = torch.cuda.current_stream().cuda_stream
        out_dev = torch.full((6,), -7.0, dtype=torch.float64, device="cuda")
        for n, seed in ((300_000, 1), (4_321, 2)):
            print(f"[no_alloc] n={n}", flush=True)
            t, ss = _state(n, seed)
            ref = (C.c_double * 6)()
            L.check(lib.sphk_diagnostics(c.handle, n, ss, ph, ref, st), "sphk_diagnostics")
            L.check(lib.sphk_diagnostics_static(c.handle, n, ss, ph, C.c_void_p(out_dev.data_ptr()), st),
                    "sphk_diagnostics_static")
            torch.cuda.synchronize()
            got = out_dev.cpu().numpy()
            assert np.array_equal(got, np.array(ref[:])), (got, ref[:])
            m = t["m"].cpu().numpy()
            assert abs(got[5] / m.sum() - 1) < 1e-12
        # a null output is rejected, not written
        rc = lib.sphk_diagnostics_static(c.handle, 10, ss, ph, None, st)
        assert rc == -1
