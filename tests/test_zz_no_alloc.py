"""phx_ch_rk4_steps_noalloc (the RK4 entry point that allocates nothing: its
workspace is a device array the library holds per GPU) against
phx_ch_rk4_steps with a caller workspace: two calls in a row, the second on a
smaller grid, so the second runs on the array the first left filled. Same
launch, same kernels: the results must be the same bits."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init

from helpers import ch_params, to_dev, to_np, stream

pytestmark = pytest.mark.gpu

CASES = [
    # n, bc, dtype, steps: the workspace of the first (3 x 130x98x66 FP64,
    # 20 MB) is larger than the second's (3 x 202x152 FP32, 0.4 MB)
    ((128, 96, 64), capi.NOFLUX, capi.F64, 7),
    ((200, 150), capi.PERIODIC, capi.F32, 5),
]


def test_noalloc_equals_workspace_entry_point(dev):
    import torch
    for n, bc, dtype, steps in CASES:
        g = capi.make_grid(n, bc=bc)
        p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
        tdt, esz = (torch.float64, 8) if dtype == capi.F64 else (torch.float32, 4)
        c0 = init.to_stored(init.ch_field(n, 4, 0.4, 0.05), g)
        a = to_dev(c0, dev, tdt)
        b = to_dev(c0, dev, tdt)
        work = torch.empty(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), dtype) // esz, dtype=tdt, device=dev)
        print(f"noalloc case {n} dtype {dtype} steps {steps}", flush=True)
        capi.call("phx_ch_rk4_steps", C.byref(g), dtype, capi.ptr(a), capi.ptr(work), C.byref(p), steps, stream())
        capi.call("phx_ch_rk4_steps_noalloc", C.byref(g), dtype, capi.ptr(b), C.byref(p), steps, stream())
        x, y = to_np(a), to_np(b)
        assert np.array_equal(x, y), (n, np.abs(x - y).max())
        assert np.abs(x - c0).max() > 0   # the state moved
