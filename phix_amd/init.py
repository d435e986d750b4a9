"""Seeded synthetic initial conditions (SplitMix64 counter stream).

SplitMix64 (Steele, Lea, Flood, OOPSLA 2014), counter form:
    x_i = mix(seed + (i+1) * 0x9E3779B97F4A7C15),   u_i = (x_i >> 11) * 2^-53
Draws are taken in physical-cell order, x fastest, then y, then z, then field
index. Sub-streams (per-run parameters, one-hot indices) use
    substream(seed, tag) = mix(seed ^ (0xD1B54A32D192ED03 * (tag + 1))).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SUB = 0xD1B54A32D192ED03
MASK = (1 << 64) - 1


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def _mix_int(z: int) -> int:
    z &= MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def substream(seed: int, tag: int) -> int:
    return _mix_int((seed ^ ((SUB * (tag + 1)) & MASK)) & MASK)


def uniform(seed: int, n: int, a: float = 0.0, b: float = 1.0, offset: int = 0) -> np.ndarray:
    """n draws of U(a, b) starting at counter `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        x = _mix(np.uint64(seed & MASK) + i * GOLDEN)
    u = (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return a + (b - a) * u


def ch_field(n, seed: int, mean: float, amp: float) -> np.ndarray:
    """Physical cells (nz, ny, nx) of c0 = mean + U(-amp, amp)."""
    shape = tuple(reversed(tuple(n) + (1,) * (3 - len(n))))
    cnt = int(np.prod(shape))
    return uniform(seed, cnt, mean - amp, mean + amp).reshape(shape)


def ac_fields(n, nfields: int, seed: int, noise: float = 1e-3) -> np.ndarray:
    """(nfields, nz, ny, nx): one-hot on i = floor(nfields*u) per cell, drawn
    from substream(seed, 1), plus U(-noise, noise) on every field from the
    main stream (field-major after the cell order)."""
    shape = tuple(reversed(tuple(n) + (1,) * (3 - len(n))))
    cnt = int(np.prod(shape))
    idx = np.floor(nfields * uniform(substream(seed, 1), cnt)).astype(np.int64)
    idx = np.minimum(idx, nfields - 1)
    out = uniform(seed, nfields * cnt, -noise, noise).reshape((nfields,) + shape)
    flat = out.reshape(nfields, cnt)
    flat[idx, np.arange(cnt)] += 1.0
    return out


def ensemble_params(nruns: int, seed: int, sizes=None):
    """Per-run (N, kappa, cbar) of the ensemble workload, from substream(seed, 2):
    N = 64 + 32*floor(15u), kappa = U(0.25, 1), cbar = U(0.3, 0.7)."""
    u = uniform(substream(seed, 2), 3 * nruns).reshape(nruns, 3)
    N = 64 + 32 * np.minimum(np.floor(15 * u[:, 0]).astype(np.int64), 14)
    kappa = 0.25 + 0.75 * u[:, 1]
    cbar = 0.3 + 0.4 * u[:, 2]
    return N, kappa, cbar


def to_stored(phys: np.ndarray, g, dtype=np.float64) -> np.ndarray:
    """Stored array (zeros in ghosts) holding `phys` on the physical cells."""
    from .capi import stored_shape, physical
    a = np.zeros(stored_shape(g), dtype=dtype)
    physical(a, g)[...] = phys
    return a
