"""Shared test helpers: seeded fields, GPU round trips through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from phix_amd import capi, init


def rand_stored(g, seed, lo=-1.0, hi=1.0):
    """Stored array with every cell (ghosts included) drawn from U(lo, hi)."""
    shp = capi.stored_shape(g)
    return init.uniform(seed, int(np.prod(shp)), lo, hi).reshape(shp)


def ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=0.5, M=1.0, dt=0.01):
    p = capi.CHParams()
    p.rho, p.ca, p.cb, p.kappa, p.M, p.dt = rho, ca, cb, kappa, M, dt
    return p


def ac_params(L=1.0, kappa=2.0, gamma=1.5, dt=0.01):
    p = capi.ACParams()
    p.L, p.kappa, p.gamma, p.dt = L, kappa, gamma, dt
    return p


def to_dev(a, dev, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def to_np(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().to("cpu").double().numpy()


def stream():
    return capi.stream_handle()


# ---------------------------------------------------------------- .field IO
def write_field(path, phys, name="c", ghost=1):
    """Reference BINARY .field: text header then physical FP64, x fastest.
    phys has numpy shape (nz, ny, nx)."""
    nz, ny, nx = phys.shape
    with open(path, "wb") as f:
        f.write(f"# PhiX ScalarField\nname    {name}\nnx {nx}  ny {ny}  nz {nz}\nghost   {ghost}\n---\n".encode())
        f.write(np.ascontiguousarray(phys, dtype="<f8").tobytes())


def write_vfield(path, phys, name="eta", ghost=1):
    """Reference BINARY .vfield; phys has shape (ncomp, nz, ny, nx)."""
    nc, nz, ny, nx = phys.shape
    with open(path, "wb") as f:
        f.write((f"# PhiX VectorField\nname         {name}\nnComponents  {nc}\nnx {nx}  ny {ny}  nz {nz}\n"
                 f"ghost        {ghost}\n---\n").encode())
        f.write(np.ascontiguousarray(phys, dtype="<f8").tobytes())


def read_field(path):
    """Returns (header dict, data) of a .field/.vfield file; data is
    (ncomp, nz, ny, nx)."""
    raw = open(path, "rb").read()
    end = raw.index(b"\n---\n") + 5
    meta = {}
    for line in raw[:end].decode().splitlines():
        toks = line.split()
        for i, t in enumerate(toks[:-1]):
            if t in ("nx", "ny", "nz", "ghost", "nComponents"):
                meta[t] = int(toks[i + 1])
    nc = meta.get("nComponents", 1)
    data = np.frombuffer(raw[end:], dtype="<f8").reshape(nc, meta["nz"], meta["ny"], meta["nx"])
    return meta, data
