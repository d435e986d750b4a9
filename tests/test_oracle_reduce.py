"""Pins the oracle's FP64 diagnostics: the mass is an exactly rounded sum
(math.fsum) and the energy is the computeFreeEnergy.py formula
(develop/Spinodal Decomposition/computeFreeEnergy.py:76-98) evaluated by numpy
on an edge/constant/wrap padded array and summed with math.fsum."""
import math

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params


def np_energy(c, d, p, mode, values=None):
    dim = c.ndim
    f_chem = p.rho * (c - p.ca) ** 2 * (p.cb - c) ** 2
    gsq = np.zeros_like(c)
    for ax in range(dim):
        pad = [(0, 0)] * dim
        pad[ax] = (1, 1)
        if mode == "wrap":
            cp = np.pad(c, pad, mode="wrap")
        elif mode == "edge":
            cp = np.pad(c, pad, mode="edge")
        else:
            cp = np.pad(c, pad, mode="constant", constant_values=0.0)
            lo = [slice(None)] * dim; lo[ax] = 0
            hi = [slice(None)] * dim; hi[ax] = -1
            cp[tuple(lo)] = values[dim - 1 - ax][0]
            cp[tuple(hi)] = values[dim - 1 - ax][1]
        sl_p = [slice(1, -1)] * dim; sl_p[ax] = slice(2, None)
        sl_m = [slice(1, -1)] * dim; sl_m[ax] = slice(0, -2)
        sl_p = [s if i == ax else slice(None) for i, s in enumerate(sl_p)]
        sl_m = [s if i == ax else slice(None) for i, s in enumerate(sl_m)]
        dd = (cp[tuple(sl_p)] - cp[tuple(sl_m)]) / (2.0 * d[dim - 1 - ax])
        gsq = gsq + dd ** 2
    return math.fsum((f_chem + 0.5 * p.kappa * gsq).ravel()) * float(np.prod(d))


@pytest.mark.parametrize("kind,mode", [(capi.PERIODIC, "wrap"), (capi.NOFLUX, "edge"), (capi.FIXED, "const")])
@pytest.mark.parametrize("n,d", [((40, 30), (0.5, 1.5)), ((12, 10, 8), (1.0, 0.7, 1.2))])
def test_mass_energy(kind, mode, n, d):
    vals = [[0.1, 0.9], [0.2, 0.8], [0.3, 0.6]]
    g = capi.make_grid(n, d=d, bc=kind, bc_value=vals)
    p = ch_params(rho=2.0, ca=0.2, cb=0.8, kappa=0.7)
    ph = init.ch_field(n, 9, 0.45, 0.3)
    c = init.to_stored(ph, g)
    m, e = O.mass_energy(g, p, c)
    assert m == math.fsum(ph.ravel()) or abs(m - math.fsum(ph.ravel())) <= 1e-15 * abs(m)
    arr = ph.reshape([n[a] for a in reversed(range(len(n)))])
    ref = np_energy(arr, d, p, mode, vals)
    assert abs(e - ref) <= 1e-13 * abs(ref)
