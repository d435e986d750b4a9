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
