"""Pins the oracle's boundary conditions (oracle/phx_oracle.c orc_bc_apply,
orc_bc_fill) against the definitions of PeriodicBC / NoFluxBC / FixedBC
(include/boundary/*.h, src/boundary/Boundary.cu:95-195) by exhaustive index
checks: every physical cell carries a unique tag, and every ghost cell must
end up holding the tag of the cell the definition names (or the constant)."""
import itertools

import numpy as np
import pytest

from phix_amd import capi
from oracle import oracle as O

P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED


def tagged(g):
    a = np.full(capi.stored_shape(g), np.nan)
    ph = capi.physical(a, g)
    ph[...] = np.arange(ph.size, dtype=np.float64).reshape(ph.shape) + 1000.0
    return a


def expected_ghost(g, a0, pos):
    """Value a ghost cell must hold after the {X, Y, Z} BC list: the last axis
    whose face the cell lies on decides (its constant, or a copy of the cell
    it maps to, which was itself filled by the earlier axes)."""
    q = list(pos)  # stored (i, j, k) in x, y, z order
    gh = g.ghost
    for a in reversed(range(g.dim)):
        p = q[a] - gh
        n = g.n[a]
        if 0 <= p < n:
            continue
        side = 0 if p < 0 else 1
        kind = g.bc[a][side]
        if kind == F:
            return g.bc_value[a][side]
        if kind == P:
            p = p + n if p < 0 else p - n
        else:
            p = 0 if p < 0 else n - 1
        q[a] = p + gh
    return a0[q[2], q[1], q[0]]


CASES = []
for dim, n in ((1, (7,)), (2, (5, 4)), (3, (4, 3, 5))):
    for ghost in (1, 2):
        kinds = [(P, P), (N, N), (F, F), (N, F), (F, N)]
        for combo in itertools.product(kinds, repeat=dim):
            CASES.append((dim, n, ghost, combo))


@pytest.mark.parametrize("dim,n,ghost,combo", CASES)
def test_bc_fill_matches_definition(dim, n, ghost, combo):
    vals = [[0.25 + a, -1.5 - a] for a in range(3)]
    g = capi.make_grid(n, ghost=ghost, bc=[list(c) for c in combo] + [[0, 0]] * (3 - dim),
                       bc_value=vals)
    a = tagged(g)
    a0 = a.copy()
    O.bc_fill(g, a)
    # physical cells untouched
    np.testing.assert_array_equal(capi.physical(a, g), capi.physical(a0, g))
    gh = ghost
    S = [g.n[ax] + 2 * gh for ax in range(3)]
    rng = [range(S[ax]) if ax < dim else range(gh, gh + 1) for ax in range(3)]
    for k in rng[2]:
        for j in rng[1]:
            for i in rng[0]:
                pos = (i, j, k)
                if all(gh <= pos[ax] < gh + g.n[ax] for ax in range(dim)):
                    continue
                assert a[k, j, i] == expected_ghost(g, a0, pos), (pos, a[k, j, i])


def test_single_bc_object_sides():
    """NoFluxBC(X, LOW) fills only the low x ghosts; FixedBC(Y, HIGH, v) only the
    high y ghosts; both over the full stored extent of the other axes."""
    g = capi.make_grid((4, 3), ghost=1)
    a = tagged(g)
    a0 = a.copy()
    O.bc_apply(g, a, 0, N, capi.LOW)
    changed = ~np.isclose(np.nan_to_num(a, nan=-7), np.nan_to_num(a0, nan=-7))
    ks, js, is_ = np.nonzero(changed)
    assert set(is_) == {0}
    np.testing.assert_array_equal(a[1, 1:4, 0], a0[1, 1:4, 1])
    b = tagged(g)
    O.bc_apply(g, b, 1, F, capi.HIGH, 3.5)
    assert np.all(b[:, 4, :] == 3.5)
    assert np.all(np.isnan(b[1, 0, :]))
