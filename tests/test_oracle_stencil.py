"""Pins the oracle's term kernels by closed forms and by the reference's own
known answers (develop/fieldOperation/test_fieldops.cu:118-291)."""
import math

import numpy as np
import pytest

from phix_amd import capi
from oracle import oracle as O

from helpers import rand_stored


def cos_mode(g, ks, phases):
    """Separable product of cos(2*pi*k*i/n + phase) on the physical cells."""
    shp = [g.n[a] for a in (2, 1, 0)]
    out = np.ones(shp)
    for a in range(g.dim):
        i = np.arange(g.n[a])
        v = np.cos(2 * math.pi * ks[a] * i / g.n[a] + phases[a])
        sh = [1, 1, 1]
        sh[2 - a] = g.n[a]
        out = out * v.reshape(sh)
    return out


@pytest.mark.parametrize("dim,n,d", [(1, (16,), (0.7,)), (2, (16, 12), (0.7, 1.3)), (3, (8, 12, 10), (0.5, 1.0, 2.0))])
def test_lap_periodic_eigenvalue(dim, n, d):
    g = capi.make_grid(n, d=d, bc=capi.PERIODIC)
    ks = [3, 2, 1][:dim]
    ph = [0.3, 1.1, -0.4][:dim]
    a = np.zeros(capi.stored_shape(g))
    capi.physical(a, g)[...] = cos_mode(g, ks, ph)
    O.bc_fill(g, a)
    rhs = np.zeros_like(a)
    O.lap_accumulate(g, rhs, a, 1.0)
    lam = sum(4.0 / d[ax] ** 2 * math.sin(math.pi * ks[ax] / n[ax]) ** 2 for ax in range(dim))
    np.testing.assert_allclose(capi.physical(rhs, g), -lam * capi.physical(a, g), atol=1e-12 * lam)
    # ghosts of rhs are untouched
    ghost = np.ones(rhs.shape, bool)
    capi.physical(ghost, g)[...] = False
    assert np.all(rhs[ghost] == 0)


def test_lap_accumulates_with_coeff():
    g = capi.make_grid((6, 5), bc=capi.PERIODIC)
    a = rand_stored(g, 3)
    O.bc_fill(g, a)
    r1 = np.zeros_like(a)
    O.lap_accumulate(g, r1, a, 1.0)
    r2 = rand_stored(g, 4)
    base = r2.copy()
    O.lap_accumulate(g, r2, a, -2.5)
    np.testing.assert_allclose(capi.physical(r2, g), capi.physical(base + -2.5 * r1, g), atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
def test_grad_central_difference_of_sine(axis):
    n, dx = (16, 12), (0.7, 1.3)
    g = capi.make_grid(n, d=dx, bc=capi.PERIODIC)
    k = 3
    i = np.arange(n[axis])
    f = np.sin(2 * math.pi * k * i / n[axis])
    a = np.zeros(capi.stored_shape(g))
    ph = capi.physical(a, g)
    ph[...] = f.reshape((1, 1, -1) if axis == 0 else (1, -1, 1))
    O.bc_fill(g, a)
    r = np.zeros_like(a)
    O.grad_accumulate(g, r, a, axis, 2.0)
    expect = 2.0 * np.cos(2 * math.pi * k * i / n[axis]) * math.sin(2 * math.pi * k / n[axis]) / dx[axis]
    np.testing.assert_allclose(capi.physical(r, g), np.broadcast_to(
        expect.reshape((1, 1, -1) if axis == 0 else (1, -1, 1)), ph.shape), atol=1e-13)


def test_fieldops_known_answers():
    """test_fieldops.cu: 4x4 mesh, f1=2, f2=3; f1*f2 = 6 (:176), coeff 2 -> 12."""
    g = capi.make_grid((4, 4), ghost=1)
    f1 = np.zeros(capi.stored_shape(g)); capi.physical(f1, g)[...] = 2.0
    f2 = np.zeros(capi.stored_shape(g)); capi.physical(f2, g)[...] = 3.0
    r = np.zeros_like(f1)
    O.mul_accumulate(g, r, f1, f2, 1.0)
    assert np.all(capi.physical(r, g) == 6.0)
    r[...] = 0
    O.mul_accumulate(g, r, f1, f2, 2.0)
    assert np.all(capi.physical(r, g) == 12.0)
    assert np.all(r[0] == 0)


def test_double_well_values():
    """f'(c) = 2*rho*(c-ca)*(c-cb)*(2c-ca-cb): roots at ca, cb, (ca+cb)/2 and
    f'(2) = 12 for f = c^2(1-c)^2; pw coefficient scales the accumulation."""
    for c in (0.0, 1.0, 0.5):
        assert O.dw(c, 1.0, 0.0, 1.0) == 0.0
    assert O.dw(2.0, 1.0, 0.0, 1.0) == 12.0
    # derivative of f(c)=rho (c-ca)^2 (c-cb)^2 by central differences
    for c in (0.13, 0.71, 1.4):
        h = 1e-6
        f = lambda x: 5.0 * (x - 0.3) ** 2 * (x - 0.7) ** 2
        assert abs(O.dw(c, 5.0, 0.3, 0.7) - (f(c + h) - f(c - h)) / (2 * h)) < 1e-6
    g = capi.make_grid((4, 4))
    s = np.zeros(capi.stored_shape(g)); capi.physical(s, g)[...] = 2.0
    r = np.zeros_like(s)
    O.pw_dw_accumulate(g, r, s, 1.0, 0.0, 1.0, 3.0)
    assert np.all(capi.physical(r, g) == 36.0)
