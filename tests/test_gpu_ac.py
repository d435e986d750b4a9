"""GPU parity of the fused multi-field Allen-Cahn RK4 kernel (config 3 family)."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ac_params, to_dev, to_np, stream

pytestmark = pytest.mark.gpu
P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED


def run_gpu(g, p, etas, nsteps, dtype, dev):
    import torch
    tdt = torch.float64 if dtype == capi.F64 else torch.float32
    ts = [to_dev(e, dev, tdt) for e in etas]
    nb = capi.lib().phx_ac_rk4_workspace_bytes(C.byref(g), dtype, len(etas))
    work = torch.zeros(nb // (8 if dtype == capi.F64 else 4), dtype=tdt, device=dev)
    arr = (C.c_void_p * len(ts))(*[capi.ptr(t) for t in ts])
    capi.call("phx_ac_rk4_steps", C.byref(g), dtype, len(ts), arr, capi.ptr(work), C.byref(p), nsteps, stream())
    return [to_np(t) for t in ts]


@pytest.mark.parametrize("nf,n,bc", [(8, (96, 80), P), (3, (70, 33), N), (5, (64, 64), [[F, N], [P, P]]),
                                      (1, (50,), P)])
def test_ac_fp64_matches_oracle(dev, nf, n, bc):
    g = capi.make_grid(n, bc=bc, bc_value=[[0.1, 0], [0, 0]] if not isinstance(bc, int) else 0.0)
    p = ac_params(L=1.0, kappa=2.0, gamma=1.5, dt=0.01)
    a = init.ac_fields(n, nf, 3)
    etas = [init.to_stored(a[f], g) for f in range(nf)]
    for chunk in range(4):
        got = run_gpu(g, p, etas, 10, capi.F64, dev)
        O.ac_rk4_steps(g, p, etas, 10)
        for f in range(nf):
            err = np.abs(capi.active(got[f], g) - capi.active(etas[f], g)).max()
            assert err <= 1e-12, (chunk, f, err)
            etas[f] = got[f].copy()  # continue from the GPU state (bit-equal when exact)


def test_ac_fp32_error(dev):
    n, nf = (128, 128), 8
    g = capi.make_grid(n, bc=P)
    p = ac_params()
    a = init.ac_fields(n, nf, 3)
    etas = [init.to_stored(a[f], g) for f in range(nf)]
    got = run_gpu(g, p, [e.copy() for e in etas], 200, capi.F32, dev)
    O.ac_rk4_steps(g, p, etas, 200)
    err = max(np.abs(capi.physical(got[f], g) - capi.physical(etas[f], g)).max() for f in range(nf))
    print(f"AC 8-field fp32 vs fp64 oracle, 128^2, 200 steps: max-abs {err:.3e}")
    assert err <= 1e-5


def _run_path(g, p, etas, nsteps, dtype, dev, two_stage):
    capi.call("phx_ac_set_two_stage", 1 if two_stage else 0)
    try:
        return run_gpu(g, p, [e.copy() for e in etas], nsteps, dtype, dev)
    finally:
        capi.call("phx_ac_set_two_stage", 1)


@pytest.mark.parametrize("dtype", [capi.F32, capi.F64])
@pytest.mark.parametrize("nf,n,bc,nsteps", [
    (8, (128, 64), P, 3),                    # full tiles, odd step count
    (8, (96, 80), P, 4),                     # even step count (buffers swapped)
    (3, (70, 33), N, 5),                     # partial tiles, zero flux
    (5, (64, 40), [[F, N], [P, P]], 2),      # fixed / zero-flux in x, periodic in y
    (2, (33, 70), [[P, P], [N, F]], 3),      # zero-flux / fixed in y
    (1, (2, 2), N, 1),                       # smallest grid: every ring cell is a ghost or an image
    (4, (3, 5), P, 2),
    (6, (130, 17), [[N, N], [P, P]], 1),
])
def test_ac_two_stage_equals_one_stage(dev, dtype, nf, n, bc, nsteps):
    """The two-stage kernel (stages 1+2 and 3+4 per launch) returns the
    one-stage kernel's results bit for bit, ghost cells included."""
    vals = [[0.1, 0.0], [0.0, 0.0]] if bc == [[F, N], [P, P]] else ([[0.0, 0.0], [0.0, -0.2]] if not isinstance(bc, int) else 0.0)
    g = capi.make_grid(n, bc=bc, bc_value=vals)
    p = ac_params(L=1.0, kappa=2.0, gamma=1.5, dt=0.01)
    a = init.ac_fields(n, nf, 7)
    etas = [init.to_stored(a[f], g) for f in range(nf)]
    one = _run_path(g, p, etas, nsteps, dtype, dev, False)
    two = _run_path(g, p, etas, nsteps, dtype, dev, True)
    print(f"{n} nf={nf} steps={nsteps} dtype={dtype}: compared", flush=True)
    for f in range(nf):
        assert np.array_equal(one[f], two[f]), (f, np.abs(one[f] - two[f]).max())
    assert max(np.abs(two[f] - etas[f]).max() for f in range(nf)) > 0   # the fields moved
