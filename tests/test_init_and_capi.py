"""CPU checks: SplitMix64 known answers, the numpy generator against the C
oracle's, and that libphix_amd.so loads and exports every symbol declared in
include/phix_amd.h. No kernel is launched here (argument errors are caught
before any launch)."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O


def test_splitmix64_known_answers():
    # Vigna's splitmix64.c from state 0: first three outputs.
    outs = []
    for i in range(1, 4):
        outs.append(init._mix_int(i * 0x9E3779B97F4A7C15))
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_numpy_generator_matches_oracle():
    for seed in (1, 2, 12345):
        np.testing.assert_array_equal(init.uniform(seed, 1000, -0.02, 0.02, offset=17),
                                      O.uniform(seed, 1000, -0.02, 0.02, offset=17))
        assert init.substream(seed, 3) == O.substream(seed, 3)


def test_ac_initial_state_one_hot():
    a = init.ac_fields((32, 16), 8, 3)
    assert a.shape == (8, 1, 16, 32)
    hot = (a > 0.5).sum(axis=0)
    assert np.all(hot == 1)
    assert np.all(np.abs(a - np.round(a)) <= 1e-3)


def test_ensemble_params_ranges():
    N, k, cb = init.ensemble_params(512, 4)
    assert set(np.unique(N)) <= set(range(64, 513, 32))
    assert len(np.unique(N)) == 15
    assert k.min() >= 0.25 and k.max() <= 1.0 and cb.min() >= 0.3 and cb.max() <= 0.7


def test_library_exports_declared_symbols():
    L = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.phx_abi_version() == 1


def test_argument_errors_are_reported():
    g = capi.make_grid((8, 8))
    g.dim = 4
    rc = capi.lib().phx_ch_rk4_steps(C.byref(g), capi.F32, C.c_void_p(1), C.c_void_p(1), None, 1, None)
    assert rc == capi.PHX_ERR_INVALID
    assert b"dim" in capi.lib().phx_last_error()
    g = capi.make_grid((8, 8), bc=[[capi.PERIODIC, capi.NOFLUX], [capi.PERIODIC, capi.PERIODIC]])
    p = capi.CHParams(1, 0, 1, 0.5, 1, 0.01)
    rc = capi.lib().phx_ch_rk4_steps(C.byref(g), capi.F32, C.c_void_p(1), C.c_void_p(1), C.byref(p), 1, None)
    assert rc == capi.PHX_ERR_INVALID and b"one side" in capi.lib().phx_last_error()
    g = capi.make_grid((8, 8), bc=capi.BC_NONE)
    rc = capi.lib().phx_ch_euler_steps(C.byref(g), capi.F64, C.c_void_p(1), C.c_void_p(2), C.byref(p), 1, None)
    assert rc == capi.PHX_ERR_UNSUPPORTED
    assert capi.lib().phx_stored_size(C.byref(capi.make_grid((128, 128)))) == 130 * 130 * 3
    with pytest.raises(capi.PhxError):
        capi.call("phx_grad_accumulate", C.byref(capi.make_grid((8, 8))), C.c_void_p(1), C.c_void_p(1), 2, 1.0, None)
    lib = capi.lib()
    # the four-step Euler tile: -1..3
    for t in (-2, 4):
        assert lib.phx_ch_euler_set_quad(t) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_euler_set_quad(-1) == 0 and lib.phx_ch_euler_set_quad(0) == 0
    # surrogate rollout with kept states: shape, NULL and `every` checks (no launch)
    fake = C.c_void_p(1)
    assert lib.phx_sur_rollout_states(fake, fake, fake, fake, None, 1, 40, 32, 4, 2, 0, None) == capi.PHX_ERR_INVALID
    assert lib.phx_sur_rollout_states(fake, fake, fake, None, None, 1, 32, 32, 4, 2, 0, None) == capi.PHX_ERR_INVALID
    assert lib.phx_sur_rollout_states(fake, fake, fake, fake, None, 1, 32, 32, 4, 3, 0, None) == capi.PHX_ERR_INVALID
    assert b"dividing" in lib.phx_last_error()
    assert lib.phx_sur_rollout_states(fake, fake, fake, fake, None, 1, 32, 32, 0, 1, 0, None) == capi.PHX_ERR_INVALID
    base = 1 << 20   # host-side range checks only: no memory is touched
    field = 32 * 32 * 4
    assert lib.phx_sur_rollout_states(fake, fake, C.c_void_p(base), C.c_void_p(base - field), None, 1, 32, 32, 4, 2, 0,
                                      None) == capi.PHX_ERR_INVALID   # states[1] is c_in
    assert b"overlaps" in lib.phx_last_error()
    assert lib.phx_sur_rollout_states(fake, fake, C.c_void_p(base), C.c_void_p(base + field), C.c_void_p(base + 2 * field),
                                      1, 32, 32, 4, 2, 0, None) == capi.PHX_ERR_INVALID   # states[1] is corr[0]


def _launch(g, dtype):
    zc, tile = C.c_int(), C.c_int()
    capi.call("phx_ch_rk4_get_launch", C.byref(g), dtype, C.byref(zc), C.byref(tile))
    return zc.value, tile.value


def test_launch_table_defaults_and_ranges():
    """The RK4 launch table is host state (no GPU needed): defaults per
    (dim, dtype), the tile ranges of 2D and 3D, the size limits of the
    temporally blocked kernels and their error codes."""
    lib = capi.lib()
    g3 = capi.make_grid((40, 36, 34), bc=capi.NOFLUX)
    assert _launch(g3, capi.F32) == (2, 11)     # stages 2-3: z-register 128x16/512 at 24 z-chunks
    assert _launch(g3, capi.F64) == (9, 17)     # stages 2-3: z-register 128x8/512 at 4 z-chunks
    g2 = capi.make_grid((1024, 1024), bc=capi.PERIODIC)
    assert _launch(g2, capi.F32) == (1, 7)      # whole-step 128x32, 1024 threads
    assert _launch(g2, capi.F64) == (1, 5)      # whole-step 32x32, 512 threads
    g2m = capi.make_grid((128, 200), bc=capi.PERIODIC)
    assert _launch(g2m, capi.F32) == (1, 5)
    g2s = capi.make_grid((7, 64), bc=capi.PERIODIC)
    assert _launch(g2s, capi.F32) == (1, 0)     # too small for the whole-step kernel
    g1 = capi.make_grid((64,), bc=capi.PERIODIC)
    assert _launch(g1, capi.F32)[1] == 0
    # ranges: 3D 0-17, 2D 0-8, 1D 0
    for t in range(18):
        assert lib.phx_ch_rk4_set_tile(C.byref(g3), capi.F32, t) == 0
    assert lib.phx_ch_rk4_set_tile(C.byref(g3), capi.F32, 18) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_rk4_set_tile(C.byref(g3), capi.F32, -1) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_rk4_set_tile(C.byref(g2), capi.F32, 9) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_rk4_set_tile(C.byref(g1), capi.F32, 1) == capi.PHX_ERR_INVALID
    # the two-stage tiles need n >= 4 per axis, the whole-step tiles n >= 8
    small3 = capi.make_grid((3, 16, 16), bc=capi.NOFLUX)
    assert lib.phx_ch_rk4_set_tile(C.byref(small3), capi.F64, 5) == capi.PHX_ERR_UNSUPPORTED
    assert lib.phx_ch_rk4_set_tile(C.byref(small3), capi.F64, 2) == 0
    assert lib.phx_ch_rk4_set_tile(C.byref(small3), capi.F64, 9) == 0   # z-register tiles: any n
    assert lib.phx_ch_rk4_set_tile(C.byref(g2s), capi.F32, 3) == capi.PHX_ERR_UNSUPPORTED
    # settings are kept per (dim, shape, dtype)
    assert lib.phx_ch_rk4_set_zchunk(C.byref(g3), capi.F32, 22) == 0
    assert _launch(g3, capi.F32) == (22, 17)
    assert _launch(g3, capi.F64) == (9, 17)
    assert lib.phx_ch_rk4_set_zchunk(C.byref(g3), capi.F32, 0) == capi.PHX_ERR_INVALID
    capi.call("phx_ch_rk4_set_tile", C.byref(g3), capi.F32, 11)
    capi.call("phx_ch_rk4_set_zchunk", C.byref(g3), capi.F32, 16)


def test_timing_record_without_launches():
    """The per-launch timing record: enabling clears it, an empty record
    reads zero launches, and kinds outside 0-9 are rejected."""
    lib = capi.lib()
    capi.call("phx_ch_timing", 1)
    for kind in range(11):
        n, ms = C.c_int(-1), C.c_double(-1.0)
        capi.call("phx_ch_timing_read", kind, C.byref(n), C.byref(ms))
        assert n.value == 0 and ms.value == 0.0
    assert lib.phx_ch_timing_read(11, None, None) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_timing_read(-1, None, None) == capi.PHX_ERR_INVALID
    capi.call("phx_ch_timing", 0)


def test_fused_paths_reject_planes_beyond_32bit_offsets():
    """A stored x-y plane of 2^31 cells or more is refused before any launch;
    below that, z-register tiles are accepted for FP64 and FP32 alike."""
    lib = capi.lib()
    big = capi.make_grid((46340, 46340, 2), bc=capi.PERIODIC)        # 46342^2 stored cells > 2^31
    p = capi.CHParams(1.0, 0.0, 1.0, 1.0, 1.0, 0.005)
    fake = C.c_void_p(0x1000)
    assert lib.phx_ch_rk4_steps(C.byref(big), capi.F32, fake, fake, C.byref(p), 1, None) == capi.PHX_ERR_UNSUPPORTED
    ok = capi.make_grid((30000, 30000, 2), bc=capi.PERIODIC)          # 3.6 GB FP32 plane: z-register ok
    assert lib.phx_ch_rk4_set_tile(C.byref(ok), capi.F32, 11) == 0
    assert lib.phx_ch_rk4_set_tile(C.byref(ok), capi.F64, 11) == capi.PHX_ERR_UNSUPPORTED   # 7.2 GB FP64 plane


def test_noalloc_rk4_rejects_oversized_grids():
    """The no-allocation RK4 entry point refuses grids whose workspace exceeds
    the array the library holds per GPU, before touching the device."""
    lib = capi.lib()
    p = capi.CHParams(1, 0, 1, 1.0, 1, 0.005)
    fake = C.c_void_p(0x1000)   # never dereferenced: the size check fails first
    big = capi.make_grid((256, 256, 256), bc=capi.NOFLUX)
    assert lib.phx_ch_rk4_workspace_bytes(C.byref(big), capi.F64) > 64 << 20
    assert lib.phx_ch_rk4_steps_noalloc(C.byref(big), capi.F64, fake, C.byref(p), 1, None) == capi.PHX_ERR_UNSUPPORTED
    assert b"per GPU" in lib.phx_last_error()
    assert lib.phx_ch_rk4_steps_noalloc(C.byref(big), capi.F64, None, C.byref(p), 1, None) == capi.PHX_ERR_INVALID


def _stage(g, dtype, cls):
    zc, tile = C.c_int(), C.c_int()
    capi.call("phx_ch_rk4_get_stage_launch", C.byref(g), dtype, cls, C.byref(zc), C.byref(tile))
    return zc.value, tile.value


def test_stage_class_launches():
    """Per stage class launches of the 3D one-stage schedule: FP64 defaults
    per class, the setters, uniform setters overriding every class, and the
    argument checks (host state only)."""
    lib = capi.lib()
    g = capi.make_grid((256, 256, 256), bc=capi.NOFLUX)
    assert [_stage(g, capi.F64, k) for k in range(3)] == [(32, 11), (64, 17), (32, 17)]
    assert [_stage(g, capi.F32, k) for k in range(3)] == [(11, 11), (11, 11), (16, 12)]
    g5 = capi.make_grid((64, 48, 40), bc=capi.NOFLUX)
    capi.call("phx_ch_rk4_set_stage_launch", C.byref(g5), capi.F32, 2, 3, 11)
    assert _stage(g5, capi.F32, 2) == (11, 3) and _stage(g5, capi.F32, 0) == (2, 11)
    assert lib.phx_ch_rk4_set_stage_launch(C.byref(g5), capi.F32, 3, 9, 16) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_rk4_set_stage_launch(C.byref(g5), capi.F32, 0, 5, 16) == capi.PHX_ERR_INVALID  # two-stage
    assert lib.phx_ch_rk4_set_stage_launch(C.byref(g5), capi.F32, 0, 18, 16) == capi.PHX_ERR_INVALID
    assert lib.phx_ch_rk4_set_stage_launch(C.byref(g5), capi.F32, 0, 9, 0) == capi.PHX_ERR_INVALID
    g2 = capi.make_grid((64, 64), bc=capi.PERIODIC)
    assert lib.phx_ch_rk4_set_stage_launch(C.byref(g2), capi.F32, 0, 0, 1) == capi.PHX_ERR_UNSUPPORTED
    capi.call("phx_ch_rk4_set_tile", C.byref(g5), capi.F32, 12)
    assert [_stage(g5, capi.F32, k)[1] for k in range(3)] == [12] * 3
