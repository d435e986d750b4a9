"""GPU parity of the 3D one-stage kernel with z-neighbours in registers.

Launch tiles 9..17 select it (TX x TY columns per workgroup of 256..1024
threads). Contract: FP64 results equal the FP64 oracle within 1e-12 max-abs
on every stored cell of the active region, and both dtypes equal the LDS-ring
one-stage kernel (tiles 0/1) bit for bit, for every BC kind, ragged shapes
(partial tiles on x and y), the smallest axes (n = 2), z-chunks of a single
plane and chunk counts that do not divide n_z."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, to_np
from test_gpu_ch_pair import launch, run_gpu

pytestmark = pytest.mark.gpu
P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED
TOL64 = 1e-12
Z_TILES = (9, 10, 11, 12, 13, 14, 15, 16, 17)
ONE_STAGE = {capi.F32: 0, capi.F64: 1}

CASES = [
    # n, d, ghost, bc, values
    ((40, 36, 34), (1.0, 1.0, 1.0), 1, N, 0.0),
    ((67, 20, 35), (1.0, 0.8, 1.1), 1, [[P, P], [N, F], [F, N]], [[0, 0], [0, 0.42], [0.38, 0]]),
    ((32, 24, 30), (1.0, 1.0, 1.0), 2, P, 0.0),
    ((37, 29, 23), (0.9, 1.2, 1.0), 1, [[F, F], [P, P], [N, N]], [[0.3, 0.5], [0, 0], [0, 0]]),
    ((5, 3, 2), (1.0, 1.0, 1.0), 1, [[N, N], [P, P], [N, F]], [[0, 0], [0, 0], [0, 0.44]]),
    ((3, 2, 2), (1.0, 1.0, 1.0), 1, P, 0.0),
    ((2, 2, 3), (1.0, 1.0, 1.0), 1, N, 0.0),
    ((270, 9, 13), (1.0, 1.0, 1.0), 1, [[N, F], [F, F], [P, P]], [[0, 0.41], [0.39, 0.43], [0, 0]]),
    # every tile inside the domain (the kernel's FULL form) for every z-register tile
    ((256, 16, 20), (1.0, 0.9, 1.1), 1, [[P, P], [N, F], [N, N]], [[0, 0], [0, 0.45], [0, 0]]),
    # FULL for the 64- and 128-wide tiles, not for the 256-wide ones; two ghost layers
    ((128, 32, 9), (1.0, 1.0, 1.0), 2, N, 0.0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_zreg_fp64_matches_oracle_and_one_stage(dev, case):
    import torch
    n, d, ghost, bc, vals = CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 31 + case, 0.4, 0.02), g)
    worst = 0.0
    for k, nsteps in enumerate((1, 3)):
        ref = c0.copy()
        O.ch_rk4_steps(g, p, ref, nsteps)
        one = run_gpu(g, capi.F64, c0, p, nsteps, dev, ONE_STAGE[capi.F64], 16)
        for tile in Z_TILES:
            zc = (1, 5, 64)[(tile + k) % 3]
            print(f"case {case} n={n} steps={nsteps} tile={tile} zc={zc}", flush=True)
            got = run_gpu(g, capi.F64, c0, p, nsteps, dev, tile, zc)
            err = np.abs(capi.active(to_np(got), g) - capi.active(ref, g)).max()
            worst = max(worst, err)
            assert err <= TOL64, (nsteps, tile, zc, err)
            assert torch.equal(got, one), (nsteps, tile, zc)
    print(f"zreg fp64 case {case}: worst max-abs vs oracle = {worst:.3e}")


@pytest.mark.parametrize("case", range(len(CASES)))
def test_zreg_fp32_equals_one_stage(dev, case):
    n, d, ghost, bc, vals = CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 41 + case, 0.4, 0.02), g)
    for nsteps in (1, 4):
        one = to_np(run_gpu(g, capi.F32, c0, p, nsteps, dev, ONE_STAGE[capi.F32], 16))
        for tile in Z_TILES:
            zc = (2, 7, 16)[(tile + nsteps) % 3]
            print(f"case {case} steps={nsteps} tile={tile} zc={zc}", flush=True)
            got = to_np(run_gpu(g, capi.F32, c0, p, nsteps, dev, tile, zc))
            assert np.array_equal(capi.active(got, g), capi.active(one, g)), (nsteps, tile, zc)


def test_zreg_full_size_256_fp32_equals_one_stage(dev):
    """The benchmark problem (256^3, zero-flux, FP32): 5 steps with each
    z-register tile equal 5 steps with the LDS-ring kernel bit for bit."""
    import torch
    n = (256, 256, 256)
    g = capi.make_grid(n, bc=N)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 2, 0.5, 0.05), g).astype(np.float32)
    one = run_gpu(g, capi.F32, c0, p, 5, dev, 0, 16)
    for tile in Z_TILES:
        print(f"256^3 tile={tile}", flush=True)
        got = run_gpu(g, capi.F32, c0, p, 5, dev, tile, 16)
        assert torch.equal(got, one), tile


@pytest.mark.parametrize("dtype", [capi.F64, capi.F32])
@pytest.mark.parametrize("case", [0, 1, 3, 7, 8])
def test_mixed_stage_class_launches_equal_uniform(dev, dtype, case):
    """Each stage class on its own one-stage launch (z-register and LDS-ring
    tiles mixed, different z-chunks) returns the uniform launch's bits."""
    import torch
    n, d, ghost, bc, vals = CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 41 + case, 0.4, 0.02), g)
    one = run_gpu(g, dtype, c0, p, 5, dev, ONE_STAGE[dtype], 16)
    for mix in ([(9, 5), (17, 8), (1, 3)], [(0, 2), (12, 16), (17, 64)], [(14, 1), (1, 7), (10, 4)]):
        with launch(g, dtype, ONE_STAGE[dtype], 16):
            for k, (tile, zc) in enumerate(mix):
                capi.call("phx_ch_rk4_set_stage_launch", C.byref(g), dtype, k, tile, zc)
            print(f"mixed launches case {case} dtype {dtype}: {mix}", flush=True)
            got = run_gpu_current(g, dtype, c0, p, 5, dev)
        assert torch.equal(got, one), mix


def run_gpu_current(g, dtype, c, p, nsteps, dev):
    import torch
    from helpers import to_dev, stream
    tdt, esz = (torch.float64, 8) if dtype == capi.F64 else (torch.float32, 4)
    tc = to_dev(c, dev, tdt)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), dtype) // esz, dtype=tdt, device=dev)
    capi.call("phx_ch_rk4_steps", C.byref(g), dtype, capi.ptr(tc), capi.ptr(work), C.byref(p), nsteps, stream())
    torch.cuda.synchronize()
    return tc
