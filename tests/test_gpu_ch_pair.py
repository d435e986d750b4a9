"""GPU parity of the temporally blocked Cahn-Hilliard kernels.

3D launch tiles 5..8 run stages (1,2) and (3,4) of each RK4 step in one kernel
each; 2D launch tiles 1..4 run a whole step in one kernel. Contract: FP64 results equal the FP64 oracle within 1e-12 max-abs on
every stored cell of the active region (physical cells and the ghost ring the
call leaves), and both dtypes equal the one-stage kernels bit for bit, for
every BC kind, ragged shapes, small axes (n = 4, the minimum) and step counts
that exercise every branch of the step schedule (odd/even counts, a one-stage
first step, the in-place last step)."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, to_dev, to_np, stream

pytestmark = pytest.mark.gpu
P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED
TOL64 = 1e-12
PAIR_TILES = (5, 6, 7, 8)
ONE_STAGE = {capi.F32: 0, capi.F64: 1}

CASES = [
    # n, d, ghost, bc, values
    ((40, 36, 34), (1.0, 1.0, 1.0), 1, N, 0.0),
    ((67, 20, 35), (1.0, 0.8, 1.1), 1, [[P, P], [N, F], [F, N]], [[0, 0], [0, 0.42], [0.38, 0]]),
    ((32, 24, 30), (1.0, 1.0, 1.0), 2, P, 0.0),
    ((37, 29, 23), (0.9, 1.2, 1.0), 1, [[F, F], [P, P], [N, N]], [[0.3, 0.5], [0, 0], [0, 0]]),
    ((4, 5, 6), (1.0, 1.0, 1.0), 1, [[N, N], [P, P], [N, F]], [[0, 0], [0, 0], [0, 0.44]]),
    ((4, 4, 4), (1.0, 1.0, 1.0), 1, P, 0.0),
    ((70, 9, 13), (1.0, 1.0, 1.0), 1, [[N, F], [F, F], [P, P]], [[0, 0.41], [0.39, 0.43], [0, 0]]),
]


class launch:
    """Select a launch (tile, z-chunk) for one shape and dtype; restore after."""

    def __init__(self, g, dtype, tile, zc):
        self.g, self.dtype, self.tile, self.zc = g, dtype, tile, zc

    def __enter__(self):
        zc, tile = C.c_int(), C.c_int()
        capi.call("phx_ch_rk4_get_launch", C.byref(self.g), self.dtype, C.byref(zc), C.byref(tile))
        self.saved = (zc.value, tile.value)
        self.classes = []
        if self.g.dim == 3:
            for k in range(3):
                capi.call("phx_ch_rk4_get_stage_launch", C.byref(self.g), self.dtype, k, C.byref(zc), C.byref(tile))
                self.classes.append((zc.value, tile.value))
        capi.call("phx_ch_rk4_set_tile", C.byref(self.g), self.dtype, self.tile)
        capi.call("phx_ch_rk4_set_zchunk", C.byref(self.g), self.dtype, self.zc)
        return self

    def __exit__(self, *exc):
        capi.call("phx_ch_rk4_set_tile", C.byref(self.g), self.dtype, self.saved[1])
        capi.call("phx_ch_rk4_set_zchunk", C.byref(self.g), self.dtype, self.saved[0])
        for k, (zc, tile) in enumerate(self.classes):
            if tile < 5 or tile >= 9:   # one-stage tiles only are per class
                capi.call("phx_ch_rk4_set_stage_launch", C.byref(self.g), self.dtype, k, tile, zc)


def run_gpu(g, dtype, c, p, nsteps, dev, tile, zc):
    import torch
    tdt, esz = (torch.float64, 8) if dtype == capi.F64 else (torch.float32, 4)
    tc = to_dev(c, dev, tdt)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), dtype) // esz, dtype=tdt, device=dev)
    with launch(g, dtype, tile, zc):
        capi.call("phx_ch_rk4_steps", C.byref(g), dtype, capi.ptr(tc), capi.ptr(work), C.byref(p), nsteps, stream())
    torch.cuda.synchronize()
    return tc


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pair_fp64_matches_oracle_and_one_stage(dev, case):
    import torch
    n, d, ghost, bc, vals = CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 3 + case, 0.4, 0.02), g)
    worst = 0.0
    for k, nsteps in enumerate((1, 2, 3, 6)):
        ref = c0.copy()
        O.ch_rk4_steps(g, p, ref, nsteps)
        one = run_gpu(g, capi.F64, c0, p, nsteps, dev, ONE_STAGE[capi.F64], 16)
        for tile in PAIR_TILES:
            zc = (3, 8, 64)[(tile + k) % 3]
            print(f"case {case} n={n} steps={nsteps} tile={tile} zc={zc}", flush=True)
            got = run_gpu(g, capi.F64, c0, p, nsteps, dev, tile, zc)
            err = np.abs(capi.active(to_np(got), g) - capi.active(ref, g)).max()
            worst = max(worst, err)
            assert err <= TOL64, (nsteps, tile, zc, err)
            assert torch.equal(got, one), (nsteps, tile, zc)
    print(f"pair fp64 case {case}: worst max-abs vs oracle = {worst:.3e}")


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pair_fp32_equals_one_stage(dev, case):
    n, d, ghost, bc, vals = CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 11 + case, 0.4, 0.02), g)
    for nsteps in (1, 4, 9):
        one = to_np(run_gpu(g, capi.F32, c0, p, nsteps, dev, ONE_STAGE[capi.F32], 16))
        for tile in PAIR_TILES:
            print(f"case {case} steps={nsteps} tile={tile}", flush=True)
            got = to_np(run_gpu(g, capi.F32, c0, p, nsteps, dev, tile, 8))
            assert np.array_equal(capi.active(got, g), capi.active(one, g)), (nsteps, tile)


def test_pair_full_size_256_fp32_equals_one_stage(dev):
    """The benchmark problem (256^3, zero-flux, FP32): 20 steps with the
    two-stage kernel equal 20 steps with the one-stage kernel bit for bit,
    and mass is conserved."""
    import torch
    n = (256, 256, 256)
    g = capi.make_grid(n, bc=N)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    one = run_gpu(g, capi.F32, c0, p, 20, dev, ONE_STAGE[capi.F32], 16)
    got = run_gpu(g, capi.F32, c0, p, 20, dev, 5, 32)
    assert torch.equal(got, one)
    m0 = capi.physical(c0, g).astype(np.float64).mean()
    m1 = capi.physical(to_np(got), g).mean()
    assert abs(m1 - m0) < 1e-6


def test_pair_tile_rejected_on_small_or_2d_grid(dev):
    g = capi.make_grid((3, 8, 8), bc=N)
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g), capi.F64, 5) == capi.PHX_ERR_UNSUPPORTED
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g), capi.F64, 9) == 0      # z-register: any n >= 2
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g), capi.F64, 18) == capi.PHX_ERR_INVALID
    g2 = capi.make_grid((64, 64), bc=P)   # 2D has tiles 0-8 only
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g2), capi.F32, 9) == capi.PHX_ERR_INVALID


def test_autotune_with_pair_candidates_keeps_parity(dev):
    """The autotuner may choose a two-stage tile; results after tuning still
    match the oracle."""
    import torch
    n = (48, 40, 36)
    g = capi.make_grid(n, bc=[[N, N], [P, P], [F, N]], bc_value=[[0, 0], [0, 0], [0.4, 0]])
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 5, 0.4, 0.02), g)
    scratch = to_dev(c0, dev)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F64) // 8, dtype=torch.float64,
                       device=dev)
    zc, ms = C.c_int(), C.c_double()
    capi.call("phx_ch_rk4_autotune", C.byref(g), capi.F64, capi.ptr(scratch), capi.ptr(work), C.byref(p), 2,
              C.byref(zc), C.byref(ms), stream())
    tile = C.c_int()
    capi.call("phx_ch_rk4_get_launch", C.byref(g), capi.F64, C.byref(zc), C.byref(tile))
    print(f"autotune: tile {tile.value} zc {zc.value} {ms.value:.4f} ms/step")
    tc = to_dev(c0, dev)
    capi.call("phx_ch_rk4_steps", C.byref(g), capi.F64, capi.ptr(tc), capi.ptr(work), C.byref(p), 5, stream())
    ref = c0.copy()
    O.ch_rk4_steps(g, p, ref, 5)
    assert np.abs(capi.active(to_np(tc), g) - capi.active(ref, g)).max() <= TOL64


def test_launch_timing_records_kinds(dev):
    """phx_ch_timing brackets launches by kind. A 5-step call with a two-stage
    tile runs 4 two-stage steps, then (1,2) and stages 3, 4 one-stage: 5 (1,2)
    launches, 4 (3,4) launches, 2 one-stage launches."""
    import torch
    n = (32, 32, 32)
    g = capi.make_grid(n, bc=N)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 9, 0.4, 0.02), g)
    tc = to_dev(c0, dev, torch.float32)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F32) // 4, dtype=torch.float32,
                       device=dev)
    with launch(g, capi.F32, 6, 16):
        capi.call("phx_ch_timing", 1)
        capi.call("phx_ch_rk4_steps", C.byref(g), capi.F32, capi.ptr(tc), capi.ptr(work), C.byref(p), 5, stream())
        counts = []
        for kind in range(3):
            k, ms = C.c_int(), C.c_double()
            capi.call("phx_ch_timing_read", kind, C.byref(k), C.byref(ms))
            counts.append(k.value)
            assert ms.value > 0.0
        capi.call("phx_ch_timing", 0)
    assert counts == [2, 5, 4], counts


# ------------------------------------------------------ 2D whole-step kernel
STEP2D_TILES = (1, 2, 3, 4, 5, 6, 7, 8)
CASES2 = [
    # n, d, ghost, bc, values
    ((64, 48), (1.0, 1.0), 1, P, 0.0),
    ((37, 29), (0.9, 1.2), 2, N, 0.0),
    ((40, 33), (1.0, 1.0), 1, [[F, N], [P, P]], [[0.45, 0], [0, 0]]),
    ((8, 8), (1.0, 1.0), 1, [[N, F], [P, P]], [[0, 0.41], [0, 0]]),
    ((130, 70), (1.0, 0.7), 1, [[F, F], [N, N]], [[0.3, 0.5], [0, 0]]),
    ((9, 150), (1.0, 1.0), 1, [[P, P], [F, N]], [[0, 0], [0.42, 0]]),
]


@pytest.mark.parametrize("case", range(len(CASES2)))
def test_step2d_fp64_matches_oracle_and_one_stage(dev, case):
    import torch
    n, d, ghost, bc, vals = CASES2[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 21 + case, 0.4, 0.02), g)
    worst = 0.0
    for nsteps in (1, 2, 3, 8):
        ref = c0.copy()
        O.ch_rk4_steps(g, p, ref, nsteps)
        one = run_gpu(g, capi.F64, c0, p, nsteps, dev, 0, 1)
        for tile in STEP2D_TILES:
            print(f"2d case {case} n={n} steps={nsteps} tile={tile}", flush=True)
            got = run_gpu(g, capi.F64, c0, p, nsteps, dev, tile, 1)
            err = np.abs(capi.active(to_np(got), g) - capi.active(ref, g)).max()
            worst = max(worst, err)
            assert err <= TOL64, (nsteps, tile, err)
            assert torch.equal(got, one), (nsteps, tile)
    print(f"step2d fp64 case {case}: worst max-abs vs oracle = {worst:.3e}")


@pytest.mark.parametrize("case", range(len(CASES2)))
def test_step2d_fp32_equals_one_stage(dev, case):
    import torch
    n, d, ghost, bc, vals = CASES2[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 31 + case, 0.4, 0.02), g)
    for nsteps in (1, 5, 10):
        one = run_gpu(g, capi.F32, c0, p, nsteps, dev, 0, 1)
        for tile in STEP2D_TILES:
            print(f"2d case {case} steps={nsteps} tile={tile}", flush=True)
            got = run_gpu(g, capi.F32, c0, p, nsteps, dev, tile, 1)
            assert torch.equal(got, one), (nsteps, tile)


def test_step2d_full_size_1024_fp32_equals_one_stage(dev):
    """The 1024^2 periodic FP32 problem: 30 steps with the whole-step kernel
    equal the one-stage kernel bit for bit; mass conserved."""
    import torch
    n = (1024, 1024)
    g = capi.make_grid(n, bc=P)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    one = run_gpu(g, capi.F32, c0, p, 30, dev, 0, 1)
    for tile in STEP2D_TILES:
        got = run_gpu(g, capi.F32, c0, p, 30, dev, tile, 1)
        assert torch.equal(got, one), tile
    m0 = capi.physical(c0, g).astype(np.float32).astype(np.float64).mean()
    assert abs(capi.physical(to_np(got), g).mean() - m0) < 1e-6


def test_step2d_tile_rejected_on_small_grid(dev):
    g = capi.make_grid((7, 64), bc=P)
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g), capi.F32, 1) == capi.PHX_ERR_UNSUPPORTED
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g), capi.F32, 9) == capi.PHX_ERR_INVALID
    g1 = capi.make_grid((64,), bc=P)
    assert capi.lib().phx_ch_rk4_set_tile(C.byref(g1), capi.F32, 1) == capi.PHX_ERR_INVALID


def test_step2d_timing_kind(dev):
    """Calls with a whole-step tile run every step as one whole-step launch,
    for even and odd step counts (no one-stage launches)."""
    import torch
    n = (64, 64)
    g = capi.make_grid(n, bc=P)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    tc = to_dev(init.to_stored(init.ch_field(n, 9, 0.4, 0.02), g), dev, torch.float32)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F32) // 4, dtype=torch.float32,
                       device=dev)
    for nsteps in (4, 5):
        with launch(g, capi.F32, 2, 1):
            capi.call("phx_ch_timing", 1)
            capi.call("phx_ch_rk4_steps", C.byref(g), capi.F32, capi.ptr(tc), capi.ptr(work), C.byref(p), nsteps,
                      stream())
            counts = []
            for kind in range(4):
                k, ms = C.c_int(), C.c_double()
                capi.call("phx_ch_timing_read", kind, C.byref(k), C.byref(ms))
                counts.append(k.value)
            capi.call("phx_ch_timing", 0)
        assert counts == [0, 0, 0, nsteps], counts


@pytest.mark.parametrize("dtype", [capi.F32, capi.F64])
@pytest.mark.parametrize("n,bc,vals,nsteps", [
    ((256, 256), P, 0.0, 7),
    ((40, 33), [[F, N], [P, P]], [[0.45, 0], [0, 0]], 4),
    ((64, 48), N, 0.0, 1),
])
@pytest.mark.parametrize("tuned", [False, True])
def test_batch_equals_per_field(dev, dtype, n, bc, vals, nsteps, tuned):
    """phx_ch_rk4_steps_batch on 5 fields leaves each one exactly as
    phx_ch_rk4_steps on that field alone (stored ghosts included), also
    after phx_ch_rk4_batch_autotune chose the batch's own launch."""
    import torch
    g = capi.make_grid(n, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, kappa=0.5, dt=0.01)
    tdt, esz = (torch.float64, 8) if dtype == capi.F64 else (torch.float32, 4)
    nb = 5
    fields = [init.to_stored(init.ch_field(n, 60 + k, 0.5, 0.05), g) for k in range(nb)]
    batch = torch.stack([to_dev(f, dev, tdt) for f in fields]).contiguous()
    wbytes = capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), dtype)
    work = torch.zeros(nb * wbytes // esz, dtype=tdt, device=dev)
    if tuned:
        scratch = batch.clone()
        ms = C.c_double()
        capi.call("phx_ch_rk4_batch_autotune", C.byref(g), dtype, nb, capi.ptr(scratch), capi.ptr(work), C.byref(p), 3,
                  C.byref(ms), stream())
        assert ms.value > 0
    capi.call("phx_ch_rk4_steps_batch", C.byref(g), dtype, nb, capi.ptr(batch), capi.ptr(work), C.byref(p), nsteps,
              stream())
    for k in range(nb):
        one = to_dev(fields[k], dev, tdt)
        w1 = torch.zeros(wbytes // esz, dtype=tdt, device=dev)
        capi.call("phx_ch_rk4_steps", C.byref(g), dtype, capi.ptr(one), capi.ptr(w1), C.byref(p), nsteps, stream())
        torch.cuda.synchronize()
        assert torch.equal(batch[k], one), k
    assert capi.lib().phx_ch_rk4_steps_batch(C.byref(g), dtype, 0, capi.ptr(batch), capi.ptr(work), C.byref(p), 1,
                                             None) == capi.PHX_ERR_INVALID
