"""ctypes binding of oracle/libphx_oracle.so -- TEST INFRASTRUCTURE ONLY.

Every function works on numpy float64 arrays in the stored (ghost-padded)
layout, shaped (sz, sy, sx). See phx_oracle.c for the reference file:line each
routine restates and for what pins it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from phix_amd.capi import Grid, CHParams, ACParams

HERE = Path(__file__).resolve().parent
LIB = HERE / "libphx_oracle.so"
_lib = None

D = C.POINTER(C.c_double)
GP = C.POINTER(Grid)
_SIG = {
    "orc_stored_size": (C.c_size_t, [GP]),
    "orc_bc_apply": (None, [GP, D, C.c_int, C.c_int, C.c_int, C.c_double]),
    "orc_bc_fill": (None, [GP, D]),
    "orc_lap_accumulate": (None, [GP, D, D, C.c_double]),
    "orc_grad_accumulate": (None, [GP, D, D, C.c_int, C.c_double]),
    "orc_mul_accumulate": (None, [GP, D, D, D, C.c_double]),
    "orc_pw_dw_accumulate": (None, [GP, D, D, C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_dw": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_ch_mu": (None, [GP, C.POINTER(CHParams), D, D]),
    "orc_ch_rhs": (None, [GP, C.POINTER(CHParams), D, D, D]),
    "orc_ch_euler_steps": (None, [GP, C.POINTER(CHParams), D, D, D, C.c_int]),
    "orc_ch_rk4_steps": (None, [GP, C.POINTER(CHParams), D, D, C.c_int]),
    "orc_ac_rk4_steps": (None, [GP, C.POINTER(ACParams), C.c_int, C.POINTER(D), D, C.c_int]),
    "orc_energy_density": (C.c_double, [GP, C.POINTER(CHParams), D, C.c_int, C.c_int, C.c_int]),
    "orc_mass_energy": (None, [GP, C.POINTER(CHParams), D, D]),
    "orc_substream": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "orc_uniform": (None, [C.c_uint64, C.c_uint64, C.c_size_t, C.c_double, C.c_double, D]),
    "orc_set_threads": (C.c_int, [C.c_int]),
    "orc_conv3x3_fwd": (None, [C.c_int] * 5 + [D, D, D, D]),
    "orc_conv3x3_dgrad": (None, [C.c_int] * 5 + [D, D, D]),
    "orc_conv3x3_wgrad": (None, [C.c_int] * 5 + [D, D, D, D]),
}


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB))
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def bc_apply(g, f, axis, kind, side, value=0.0):
    lib().orc_bc_apply(C.byref(g), _p(f), axis, kind, side, value)


def bc_fill(g, f):
    lib().orc_bc_fill(C.byref(g), _p(f))


def lap_accumulate(g, rhs, src, coeff):
    lib().orc_lap_accumulate(C.byref(g), _p(rhs), _p(src), coeff)


def grad_accumulate(g, rhs, src, axis, coeff):
    lib().orc_grad_accumulate(C.byref(g), _p(rhs), _p(src), axis, coeff)


def mul_accumulate(g, rhs, s1, s2, coeff):
    lib().orc_mul_accumulate(C.byref(g), _p(rhs), _p(s1), _p(s2), coeff)


def pw_dw_accumulate(g, rhs, src, rho, ca, cb, coeff):
    lib().orc_pw_dw_accumulate(C.byref(g), _p(rhs), _p(src), rho, ca, cb, coeff)


def dw(c, rho, ca, cb):
    return lib().orc_dw(c, rho, ca, cb)


def ch_mu(g, p, c, mu):
    lib().orc_ch_mu(C.byref(g), C.byref(p), _p(c), _p(mu))


def ch_rhs(g, p, u):
    mu = np.zeros_like(u)
    rhs = np.zeros_like(u)
    lib().orc_ch_rhs(C.byref(g), C.byref(p), _p(u), _p(mu), _p(rhs))
    return rhs


def ch_euler_steps(g, p, c, mu, nsteps):
    scratch = np.zeros_like(c)
    lib().orc_ch_euler_steps(C.byref(g), C.byref(p), _p(c), _p(mu), _p(scratch), nsteps)


def ch_rk4_steps(g, p, c, nsteps, work=None):
    if work is None:
        work = np.zeros((6,) + c.shape)
    lib().orc_ch_rk4_steps(C.byref(g), C.byref(p), _p(c), _p(work), nsteps)
    return work


def ac_rk4_steps(g, p, etas, nsteps, work=None):
    nf = len(etas)
    if work is None:
        work = np.zeros((5 * nf,) + etas[0].shape)
    arr = (D * nf)(*[_p(e) for e in etas])
    lib().orc_ac_rk4_steps(C.byref(g), C.byref(p), nf, arr, _p(work), nsteps)
    return work


def mass_energy(g, p, c):
    out = np.zeros(2)
    lib().orc_mass_energy(C.byref(g), C.byref(p), _p(c), _p(out))
    return float(out[0]), float(out[1])


def energy_density(g, p, c, i, j, k):
    return lib().orc_energy_density(C.byref(g), C.byref(p), _p(c), i, j, k)


def set_threads(n: int) -> int:
    """Threads of the per-cell passes (1 = serial, as the reference's CPU
    path); returns the count now in effect. Results do not depend on it."""
    return lib().orc_set_threads(int(n))


def uniform(seed, n, a, b, offset=0):
    out = np.zeros(n)
    lib().orc_uniform(seed, offset, n, a, b, _p(out))
    return out


def substream(seed, tag):
    return int(lib().orc_substream(seed, tag))


def conv3x3_fwd(x, w, b=None):
    N, Ci, H, W = x.shape
    Co = w.shape[0]
    y = np.zeros((N, Co, H, W))
    x = np.ascontiguousarray(x, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    bp = _p(np.ascontiguousarray(b, np.float64)) if b is not None else None
    lib().orc_conv3x3_fwd(N, Ci, Co, H, W, _p(x), _p(w), bp, _p(y))
    return y


def conv3x3_dgrad(dy, w):
    N, Co, H, W = dy.shape
    Ci = w.shape[1]
    dx = np.zeros((N, Ci, H, W))
    lib().orc_conv3x3_dgrad(N, Ci, Co, H, W, _p(np.ascontiguousarray(dy, np.float64)),
                            _p(np.ascontiguousarray(w, np.float64)), _p(dx))
    return dx


def conv3x3_wgrad(x, dy):
    N, Ci, H, W = x.shape
    Co = dy.shape[1]
    dw = np.zeros((Co, Ci, 3, 3))
    db = np.zeros(Co)
    lib().orc_conv3x3_wgrad(N, Ci, Co, H, W, _p(np.ascontiguousarray(x, np.float64)),
                            _p(np.ascontiguousarray(dy, np.float64)), _p(dw), _p(db))
    return dw, db
