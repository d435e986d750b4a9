"""GPU parity of the boundary and per-term kernels against the FP64 oracle,
through the C ABI. These keep the reference's operation order and must agree
bit for bit."""
import ctypes as C
import itertools

import numpy as np
import pytest

from phix_amd import capi
from oracle import oracle as O

from helpers import rand_stored, to_dev, to_np, stream

pytestmark = pytest.mark.gpu
P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED


def grids():
    out = []
    for n in ((9,), (7, 6), (5, 4, 6)):
        dim = len(n)
        for ghost in (1, 2):
            for combo in [((P, P),) * dim, ((N, N),) * dim, ((F, N), (N, F), (P, P))[:dim],
                          ((P, P), (F, F), (N, F))[:dim]]:
                vals = [[0.25 + a, -0.5 - a] for a in range(3)]
                bc = [list(c) for c in combo] + [[0, 0]] * (3 - dim)
                out.append(capi.make_grid(n, d=(0.7, 1.1, 1.3)[:dim], ghost=ghost, bc=bc, bc_value=vals))
    return out


@pytest.mark.parametrize("gi", range(len(grids())))
@pytest.mark.parametrize("dtype", [capi.F64, capi.F32])
def test_bc_fill_and_apply(dev, gi, dtype):
    import torch
    g = grids()[gi]
    a = rand_stored(g, 100 + gi)
    if dtype == capi.F32:
        a = a.astype(np.float32).astype(np.float64)
    tdt = torch.float64 if dtype == capi.F64 else torch.float32
    t = to_dev(a, dev, tdt)
    capi.call("phx_bc_fill", C.byref(g), dtype, capi.ptr(t), stream())
    O.bc_fill(g, a)
    np.testing.assert_array_equal(to_np(t), a)
    # single objects: NoFlux low on the last axis, Fixed high on x
    b = rand_stored(g, 200 + gi)
    if dtype == capi.F32:
        b = b.astype(np.float32).astype(np.float64)
    t = to_dev(b, dev, tdt)
    ax = g.dim - 1
    capi.call("phx_bc_apply", C.byref(g), dtype, capi.ptr(t), ax, N, capi.LOW, 0.0, stream())
    capi.call("phx_bc_apply", C.byref(g), dtype, capi.ptr(t), 0, F, capi.HIGH, 0.375, stream())
    O.bc_apply(g, b, ax, N, capi.LOW)
    O.bc_apply(g, b, 0, F, capi.HIGH, 0.375)
    np.testing.assert_array_equal(to_np(t), b)


@pytest.mark.parametrize("gi", range(len(grids())))
def test_term_kernels_bit_exact(dev, gi):
    g = grids()[gi]
    src = rand_stored(g, 300 + gi)
    s2 = rand_stored(g, 400 + gi)
    rhs = rand_stored(g, 500 + gi)
    ts, ts2, tr = to_dev(src, dev), to_dev(s2, dev), to_dev(rhs, dev)
    capi.call("phx_lap_accumulate", C.byref(g), capi.ptr(tr), capi.ptr(ts), -0.37, stream())
    O.lap_accumulate(g, rhs, src, -0.37)
    for ax in range(g.dim):
        capi.call("phx_grad_accumulate", C.byref(g), capi.ptr(tr), capi.ptr(ts), ax, 1.7, stream())
        O.grad_accumulate(g, rhs, src, ax, 1.7)
    capi.call("phx_mul_accumulate", C.byref(g), capi.ptr(tr), capi.ptr(ts), capi.ptr(ts2), 0.3, stream())
    O.mul_accumulate(g, rhs, src, s2, 0.3)
    capi.call("phx_pw_doublewell_accumulate", C.byref(g), capi.ptr(tr), capi.ptr(ts), 5.0, 0.3, 0.7, 1.0, stream())
    O.pw_dw_accumulate(g, rhs, src, 5.0, 0.3, 0.7, 1.0)
    np.testing.assert_array_equal(to_np(tr), rhs)


def test_flat_kernels(dev):
    import torch
    n = 100003
    rng = np.random.default_rng(0)
    a, b, k1, k2, k3, k4 = (rng.standard_normal(n) for _ in range(6))
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    tk = [to_dev(x, dev) for x in (k1, k2, k3, k4)]
    capi.call("phx_axpy_f64", capi.ptr(ta), capi.ptr(tb), 0.25, n, stream())
    a = a + 0.25 * b
    np.testing.assert_array_equal(to_np(ta), a)
    capi.call("phx_rk4_tmp_f64", capi.ptr(tb), capi.ptr(ta), capi.ptr(tk[0]), 0.005, n, stream())
    b = a + 0.005 * k1
    np.testing.assert_array_equal(to_np(tb), b)
    capi.call("phx_rk4_update_f64", capi.ptr(ta), *[capi.ptr(t) for t in tk], 0.01 / 6, n, stream())
    a = a + (0.01 / 6) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    np.testing.assert_array_equal(to_np(ta), a)
    capi.call("phx_copy_f64", capi.ptr(tb), capi.ptr(ta), n, stream())
    np.testing.assert_array_equal(to_np(tb), a)
    capi.call("phx_fill_f64", capi.ptr(tb), -3.5, n, stream())
    assert np.all(to_np(tb) == -3.5)
