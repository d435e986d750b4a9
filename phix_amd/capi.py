"""ctypes binding of libphix_amd.so (include/phix_amd.h).

This is plumbing for the tests and the benchmark: it loads the in-tree library
and declares every entry point. There is no fallback: if the library is
missing the import of `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libphix_amd.so"

PHX_OK, PHX_ERR_INVALID, PHX_ERR_UNSUPPORTED, PHX_ERR_HIP, PHX_ERR_NOMEM = 0, -1, -2, -3, -4
F64, F32 = 0, 1
BC_NONE, PERIODIC, NOFLUX, FIXED = 0, 1, 2, 3
LOW, HIGH, BOTH = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("n", C.c_int * 3),
        ("ghost", C.c_int),
        ("d", C.c_double * 3),
        ("bc", (C.c_int * 2) * 3),
        ("bc_value", (C.c_double * 2) * 3),
    ]


class CHParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("rho", "ca", "cb", "kappa", "M", "dt")]


class ACParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("L", "kappa", "gamma", "dt")]


class GfaParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("dt", "M_eta", "eps_sq", "beta_sq", "w_phi", "w_eta", "w_ex", "T", "T_g",
                                          "alpha", "V_m_L", "V_m_S", "D_Am")]


def make_grid(n, d=(1.0, 1.0, 1.0), ghost=1, bc=PERIODIC, bc_value=0.0) -> Grid:
    """n: tuple of active sizes (len = dim). bc: one kind for every face, or a
    list per axis of (lo, hi) kinds. bc_value likewise."""
    dim = len(n)
    g = Grid()
    g.dim = dim
    for a in range(3):
        g.n[a] = int(n[a]) if a < dim else 1
        g.d[a] = float(d[a]) if a < len(d) else 1.0
    g.ghost = ghost
    for a in range(3):
        for s in range(2):
            if a >= dim:
                g.bc[a][s] = BC_NONE
                g.bc_value[a][s] = 0.0
                continue
            if isinstance(bc, int):
                g.bc[a][s] = bc
            else:
                g.bc[a][s] = bc[a][s]
            if isinstance(bc_value, (int, float)):
                g.bc_value[a][s] = float(bc_value)
            else:
                g.bc_value[a][s] = float(bc_value[a][s])
    return g


def stored_shape(g: Grid):
    """numpy shape (sz, sy, sx) of one stored field."""
    return tuple(g.n[a] + 2 * g.ghost for a in (2, 1, 0))


def physical(arr, g: Grid):
    """View of the physical cells of a stored array shaped stored_shape(g)."""
    gh = g.ghost
    return arr[gh:gh + g.n[2], gh:gh + g.n[1], gh:gh + g.n[0]]


def active(arr, g: Grid):
    """Stored cells whose coordinates on the unused axes are physical: the
    physical cells plus the ghost ring of the active axes."""
    gh = g.ghost
    sl = []
    for a in (2, 1, 0):
        if a < g.dim:
            sl.append(slice(None))
        else:
            sl.append(slice(gh, gh + 1))
    return arr[tuple(sl)]


class PhxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phx error {code}: {msg}")
        self.code = code


_lib = None

VP = C.c_void_p
GP = C.POINTER(Grid)
SIG = {
    "phx_abi_version": (C.c_int, []),
    "phx_last_error": (C.c_char_p, []),
    "phx_stored_size": (C.c_size_t, [GP]),
    "phx_stream_sync": (C.c_int, [VP]),
    "phx_bc_apply": (C.c_int, [GP, C.c_int, VP, C.c_int, C.c_int, C.c_int, C.c_double, VP]),
    "phx_bc_fill": (C.c_int, [GP, C.c_int, VP, VP]),
    "phx_lap_accumulate": (C.c_int, [GP, VP, VP, C.c_double, VP]),
    "phx_grad_accumulate": (C.c_int, [GP, VP, VP, C.c_int, C.c_double, VP]),
    "phx_mul_accumulate": (C.c_int, [GP, VP, VP, VP, C.c_double, VP]),
    "phx_pw_doublewell_accumulate": (C.c_int, [GP, VP, VP, C.c_double, C.c_double, C.c_double, C.c_double, VP]),
    "phx_fill_f64": (C.c_int, [VP, C.c_double, C.c_size_t, VP]),
    "phx_copy_f64": (C.c_int, [VP, VP, C.c_size_t, VP]),
    "phx_axpy_f64": (C.c_int, [VP, VP, C.c_double, C.c_size_t, VP]),
    "phx_rk4_tmp_f64": (C.c_int, [VP, VP, VP, C.c_double, C.c_size_t, VP]),
    "phx_rk4_update_f64": (C.c_int, [VP, VP, VP, VP, VP, C.c_double, C.c_size_t, VP]),
    "phx_convert_f64_to_f32": (C.c_int, [VP, VP, C.c_size_t, VP]),
    "phx_convert_f32_to_f64": (C.c_int, [VP, VP, C.c_size_t, VP]),
    "phx_ch_euler_set_fused": (C.c_int, [C.c_int]),
    "phx_ch_euler_set_quad": (C.c_int, [C.c_int]),
    "phx_ch_euler_steps": (C.c_int, [GP, C.c_int, VP, VP, C.POINTER(CHParams), C.c_int, VP]),
    "phx_ch_rk4_workspace_bytes": (C.c_size_t, [GP, C.c_int]),
    "phx_gfa_workspace_bytes": (C.c_size_t, [GP]),
    "phx_gfa_euler_steps": (C.c_int, [GP, VP, VP, VP, VP, C.POINTER(GfaParams), C.c_int, VP]),
    "phx_ch_rk4_steps": (C.c_int, [GP, C.c_int, VP, VP, C.POINTER(CHParams), C.c_int, VP]),
    "phx_ch_rk4_steps_batch": (C.c_int, [GP, C.c_int, C.c_int, VP, VP, C.POINTER(CHParams), C.c_int, VP]),
    "phx_ch_rk4_batch_autotune": (C.c_int, [GP, C.c_int, C.c_int, VP, VP, C.POINTER(CHParams), C.c_int,
                                            C.POINTER(C.c_double), VP]),
    "phx_ch_rk4_steps_noalloc": (C.c_int, [GP, C.c_int, VP, C.POINTER(CHParams), C.c_int, VP]),
    "phx_ch_rk4_set_zchunk": (C.c_int, [GP, C.c_int, C.c_int]),
    "phx_ch_set_xcd_remap": (C.c_int, [C.c_int]),
    "phx_ch_set_fp64_exact": (C.c_int, [C.c_int]),
    "phx_ch_get_fp64_exact": (C.c_int, []),
    "phx_ch_rk4_get_launch": (C.c_int, [GP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "phx_ch_rk4_set_tile": (C.c_int, [GP, C.c_int, C.c_int]),
    "phx_ch_rk4_set_stage_launch": (C.c_int, [GP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "phx_ch_rk4_get_stage_launch": (C.c_int, [GP, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "phx_ch_rk4_autotune": (C.c_int, [GP, C.c_int, VP, VP, C.POINTER(CHParams), C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), VP]),
    "phx_ch_timing": (C.c_int, [C.c_int]),
    "phx_ch_timing_read": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "phx_reduce_workspace_bytes": (C.c_size_t, [GP]),
    "phx_mass_energy": (C.c_int, [GP, C.c_int, VP, C.POINTER(CHParams), VP, VP, VP]),
    "phx_mass_energy_batch": (C.c_int, [GP, C.c_int, C.c_int, VP, C.POINTER(CHParams), VP, VP, VP]),
    "phx_ac_rk4_workspace_bytes": (C.c_size_t, [GP, C.c_int, C.c_int]),
    "phx_ac_rk4_steps": (C.c_int, [GP, C.c_int, C.c_int, C.POINTER(VP), VP, C.POINTER(ACParams), C.c_int, VP]),
    "phx_ac_set_two_stage": (C.c_int, [C.c_int]),
    "phx_ensemble_create": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(VP)]),
    "phx_ensemble_destroy": (C.c_int, [VP]),
    "phx_ensemble_cells": (C.c_longlong, [VP]),
    "phx_ensemble_offset": (C.c_longlong, [VP, C.c_int]),
    "phx_ensemble_rk4_steps": (C.c_int, [VP, VP, VP, C.c_double, C.c_int, VP]),
    "phx_ensemble_graph_steps": (C.c_int, [VP, VP, VP, C.c_double, C.c_int, C.c_int, VP]),
    "phx_ensemble_mass": (C.c_int, [VP, VP, VP, VP]),
    "phx_ensemble_set_tiling": (C.c_int, [VP, C.c_int, C.c_int]),
    "phx_ensemble_get_tiling": (C.c_int, [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]),
    "phx_ensemble_set_two_stage": (C.c_int, [VP, C.c_int]),
    "phx_ensemble_get_two_stage": (C.c_int, [VP]),
    "phx_ensemble_autotune": (C.c_int, [VP, VP, VP, C.c_double, C.c_int, C.POINTER(C.c_double), VP]),
    "phx_sur_param_count": (C.c_int, []),
    "phx_sur_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "phx_sur_forward": (C.c_int, [VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, VP]),
    "phx_sur_rollout": (C.c_int, [VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, VP]),
    "phx_sur_rollout_states": (C.c_int, [VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         VP]),
    "phx_sur_backward": (C.c_int, [VP, VP, VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, VP]),
    "phx_sur_adam": (C.c_int, [VP, VP, VP, VP, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, VP]),
    "phx_conv3x3_ws_bytes": (C.c_size_t, []),
    "phx_conv3x3_bf16_fwd": (C.c_int, [VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, VP]),
    "phx_conv3x3_bf16_dgrad": (C.c_int, [VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, VP]),
    "phx_conv3x3_bf16_wgrad": (C.c_int, [VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int, VP]),
}


def lib():
    """Loads the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m phix_amd.build` (no fallback exists)")
    L = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIG.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int):
    if rc != PHX_OK:
        raise PhxError(rc, lib().phx_last_error().decode())
    return rc


def call(name, *args):
    """Calls an int-returning entry point and raises PhxError on failure."""
    return check(getattr(lib(), name)(*args))


def ptr(t) -> int:
    """Device (or host) address of a torch tensor or numpy array."""
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(stream=None):
    """hipStream_t of a torch stream (None -> torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def declared_symbols(header: Path | None = None) -> list[str]:
    """Names of the functions declared in include/phix_amd.h."""
    import re
    h = header or (PKG.parent / "include" / "phix_amd.h")
    txt = h.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", txt)))
