"""FP64 numpy restatement of the generic DSL's composite forms and of the
CH+AC program -- TEST INFRASTRUCTURE ONLY (the checker of the device and host
paths; never imported by the library).

Grids are numpy (ny, nx) arrays of physical cells. A ghost layer of one cell
is added by `pad` with the reference's BC fills, x faces first, then y faces
over the full padded extent (src/boundary/Boundary.cu:204-330):
  periodic: ghost = the cell one period away;
  noflux:   ghost = the face cell.
Stencils (src/equation/Equation.cu:40-124, 141-143, 212-237):
  gx  = (s[i+1] - s[i-1]) * (0.5 / dx)
  lap = (s[i+1] - 2 s[i] + s[i-1]) * (1/dx^2) + (same along y) * (1/dy^2)
A composite term -- grad(expr, axis, bcs), lap(expr, bcs), div(vector expr,
bcs) -- materialises its expression on the physical cells, applies the BCs
and then takes the stencil (src/equation/Equation.cu:537-743,
include/equation/FieldOps.inl:136-250); Term * Term multiplies the two terms'
values cell by cell.

What pins each function (tests/test_oracle_dsl.py): gx, gy and lap against
their closed forms on Fourier modes (periodic) and on the half-integer cosine
modes that are the noflux eigenfunctions, every form of composite_forms
against closed forms on such modes, and chac_steps through the free energy it
descends (its pointwise forces as finite-difference derivatives of one
density, an independent scalar integration for uniform fields, exact mode
amplification of the linearised c equation, and conservation of sum c).
"""
from __future__ import annotations

import numpy as np


def pad(f, mode):
    """One ghost layer per side; mode "periodic" or "noflux" on both axes."""
    ny, nx = f.shape
    s = np.zeros((ny + 2, nx + 2))
    s[1:-1, 1:-1] = f
    if mode == "periodic":
        s[:, 0], s[:, -1] = s[:, nx], s[:, 1]
        s[0, :], s[-1, :] = s[ny, :], s[1, :]
    elif mode == "noflux":
        s[:, 0], s[:, -1] = s[:, 1], s[:, nx]
        s[0, :], s[-1, :] = s[1, :], s[ny, :]
    else:
        raise ValueError(mode)
    return s


def gx(f, dx, mode):
    s = pad(f, mode)
    return (s[1:-1, 2:] - s[1:-1, :-2]) * (0.5 / dx)


def gy(f, dy, mode):
    s = pad(f, mode)
    return (s[2:, 1:-1] - s[:-2, 1:-1]) * (0.5 / dy)


def lap(f, dx, dy, mode):
    s = pad(f, mode)
    c = s[1:-1, 1:-1]
    return (s[1:-1, 2:] - 2.0 * c + s[1:-1, :-2]) * (1.0 / (dx * dx)) + \
        (s[2:, 1:-1] - 2.0 * c + s[:-2, 1:-1]) * (1.0 / (dy * dy))


def composite_forms(a, b, v0, v1, dx, dy, mode):
    """The forms tests/cpp/composite.cpp evaluates through the drop-in DSL:
      gradx_sq   grad(pw(a, u^2), 0, bcs)
      grady_ab   grad(a * b, 1, bcs)
      lap_ab     lap(a * b, bcs)
      divgrad_ab div(grad(a * b, bcs), bcs)
      ab_lapb    (a * b) * lap(b)           (b's ghosts BC-filled beforehand)
      mixed      2 grad(a, 0) + div(v) - a * b
    Returns {name: (ny, nx)}."""
    ab = a * b
    return {
        "gradx_sq": gx(a * a, dx, mode),
        "grady_ab": gy(ab, dy, mode),
        "lap_ab": lap(ab, dx, dy, mode),
        "divgrad_ab": gx(gx(ab, dx, mode), dx, mode) + gy(gy(ab, dy, mode), dy, mode),
        "ab_lapb": ab * lap(b, dx, dy, mode),
        "mixed": 2.0 * gx(a, dx, mode) + (gx(v0, dx, mode) + gy(v1, dy, mode)) - ab,
    }


def composite_inputs(nx=24, ny=20, dx=0.5, dy=0.25):
    """The fields of tests/cpp/composite.cpp at cell centres (i + 1/2) d:
    a = 1.5 + sin(x) cos(2y), b = cos(0.7x - 1.3y), v = (x sin y, exp(-0.1x) y)."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    return (1.5 + np.sin(X) * np.cos(2.0 * Y), np.cos(0.7 * X - 1.3 * Y), X * np.sin(Y), np.exp(-0.1 * X) * Y)


def chac_h(eta):
    """h(eta) = eta^3 (6 eta^2 - 15 eta + 10) (CH+AC_double-well.cu:100-101)."""
    return eta ** 3 * (6.0 * eta ** 2 - 15.0 * eta + 10.0)


def chac_steps(c, eta, k, dt, nsteps, mode="periodic", dx=1.0, dy=1.0):
    """The CH_AC_2D program's multi-step Euler steps
    (applications/solvers/Cahn-Hillard+Allen-Cahn_double-well/2D/
    CH+AC_double-well.cu:95-136, Solver.cu:317-338), in its order:
      1. mu   = 2 rho^2 (c - ca) + 2 rho^2 (ca - cb) h(eta) - kappa_c lap(c)   STEADY
      2. c   += dt M lap(mu)                                                   TRANSIENT
      3. eta += dt [-L (30 rho^2 eta^2 (1 - eta)^2 (2c - ca - cb)(ca - cb)
                        + 2 w eta (1 - eta)(1 - 2 eta)) + L kappa_eta lap(eta)]  TRANSIENT
    Step 3 reads the updated c. Returns (c, eta, mu)."""
    rho, ca, cb = k["rho"], k["ca"], k["cb"]
    c, eta = np.array(c, dtype=np.float64, copy=True), np.array(eta, dtype=np.float64, copy=True)
    mu = None
    for _ in range(nsteps):
        mu = 2.0 * rho * rho * (c - ca) + 2.0 * rho * rho * (ca - cb) * chac_h(eta) - k["kappa_c"] * lap(c, dx, dy, mode)
        c = c + dt * (k["M"] * lap(mu, dx, dy, mode))
        bulk = 30.0 * rho * rho * eta ** 2 * (1.0 - eta) ** 2 * (2.0 * c - ca - cb) * (ca - cb)
        well = 2.0 * k["w"] * eta * (1.0 - eta) * (1.0 - 2.0 * eta)
        eta = eta + dt * (-k["L"] * (bulk + well) + k["L"] * k["kappa_eta"] * lap(eta, dx, dy, mode))
    return c, eta, mu
