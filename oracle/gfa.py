"""FP64 numpy restatement of the Fe-B glass-formation model and its multi-step
Euler step -- TEST INFRASTRUCTURE ONLY.

Restates applications/solvers/glass_formation/2D/GFA.cu of the reference
expression by expression:

    Gibbs energies, L, dG_f(Fe3B)       GFA.cu:19-52
    Inden f(tau)                        GFA.cu:53-62
    h, h', g'                           GFA.cu:64-66
    dG_L/dc, G_L, f_L, f_S, df_Am       GFA.cu:68-85
    Arrhenius mobilities                GFA.cu:117-121
    the six multi-steps of one step     GFA.cu:143-201 (Solver.cu:317-338 order)

Fields are 2D float64 arrays of the physical cells, shaped (ny, nx); ghost
cells (one layer) are rebuilt from the BCs where a stencil reads them, which
equals the reference's BC application on the stored arrays at each multi-step
(PeriodicBC wraps, NoFluxBC mirrors the first interior cell).

Pinned by tests/test_oracle_gfa.py: breakpoint continuity of every SGTE piece
(value and slope) except the Fe pieces at 1811 K, whose jump is exactly the
reference's 773597/T in place of SGTE's 77359/T; the Inden decimals against
their closed form (p = 0.6); mu, dF/dphi and dF/deta as partial derivatives of
one free-energy density; dG_L/dc as the derivative of G_L; the mobility's pure
phase limits; lap and grad against the pinned C oracle; second-order
convergence of the c operator to div(D grad mu).
"""
from __future__ import annotations

import math

import numpy as np

R_GAS = 8.314


def G_B_liquid(T):
    if T <= 500.0:
        return (40723.275 + 86.843839 * T - 15.6641 * T * math.log(T) - 6.864515e-3 * T * T
                + 0.618878e-6 * T * T * T + 370843.0 / T)
    if T <= 2348.0:
        return (41119.703 + 82.101722 * T - 14.9827763 * T * math.log(T) - 7.095669e-3 * T * T
                + 0.507347e-6 * T * T * T + 335484.0 / T)
    return 28842.012 + 200.94731 * T - 31.4 * T * math.log(T)


def G_Fe_liquid(T):
    if T <= 1811.0:
        T2 = T * T
        T3 = T2 * T
        T7 = T3 * T3 * T
        return (13265.87 + 117.57557 * T - 23.5143 * T * math.log(T) - 4.39752e-3 * T2 - 0.058927e-6 * T3
                + 773597.0 / T - 367.516e-23 * T7)
    return -10838.83 + 291.302 * T - 46.0 * T * math.log(T)


def L_BFe_liquid(T):
    return -122861.0 + 14.59 * T


def G_Fe_bcc(T):
    if T <= 1811.0:
        return (1225.7 + 124.134 * T - 23.5143 * T * math.log(T) - 4.39752e-3 * T * T - 0.058927e-6 * T * T * T
                + 773597.0 / T)
    T3 = T * T * T
    T9 = T3 * T3 * T3
    return -25383.581 + 299.31255 * T - 46.0 * T * math.log(T) + 2296.03e28 / T9


def G_B_beta(T):
    if T <= 1100.0:
        return (-7735.284 + 107.111864 * T - 15.6641 * T * math.log(T) - 6.864515e-3 * T * T
                + 0.618878e-6 * T * T * T + 370843.0 / T)
    if T <= 2348.0:
        return (-16649.474 + 184.801744 * T - 26.6047 * T * math.log(T) - 0.79809e-3 * T * T
                - 0.02556e-6 * T * T * T + 1748270.0 / T)
    if T <= 3000.0:
        return (-36667.582 + 231.336244 * T - 31.5957527 * T * math.log(T) - 1.59488e-3 * T * T
                + 0.134719e-6 * T * T * T + 11205883.0 / T)
    return -21530.653 + 222.396264 * T - 31.4 * T * math.log(T)


def deltaG_Fe3B(T):
    return -77749.0 + 2.59 * T


def f_tau(tau):
    if tau <= 1.0:
        t3 = tau * tau * tau
        t9 = t3 * t3 * t3
        t15 = t9 * t3 * t3
        return 1.0 - 9.9167285e-1 / tau - 1.11737779e-1 * t3 - 4.96612349e-3 * t9 - 1.11737779e-3 * t15
    inv = 1.0 / tau
    i5 = inv * inv * inv * inv * inv
    i10 = i5 * i5
    i15 = i10 * i5
    i25 = i15 * i10
    return -1.05443689e-1 * i5 - 3.34741816e-3 * i15 - 7.02957924e-4 * i25


def h(x):
    return x * x * x * (10.0 - 15.0 * x + 6.0 * x * x)


def h_prime(x):
    return 30.0 * x * x * (1.0 - x) * (1.0 - x)


def g_prime(x):
    return 2.0 * x * (1.0 - x) * (1.0 - 2.0 * x)


def _clamp(c):
    return np.maximum(1e-12, np.minimum(1.0 - 1e-12, c))


def dGL_dc(c, T):
    cs = _clamp(c)
    return (G_B_liquid(T) - G_Fe_liquid(T)) + R_GAS * T * np.log(cs / (1.0 - cs)) + (1.0 - 2.0 * cs) * L_BFe_liquid(T)


def G_L(c, T):
    cs = _clamp(c)
    return (cs * G_B_liquid(T) + (1.0 - cs) * G_Fe_liquid(T)
            + R_GAS * T * (cs * np.log(cs) + (1.0 - cs) * np.log(1.0 - cs)) + cs * (1.0 - cs) * L_BFe_liquid(T))


def f_L(c, T, VmL):
    return G_L(c, T) / VmL


def f_S(T, VmS):
    return (3.0 * G_Fe_bcc(T) + G_B_beta(T) + deltaG_Fe3B(T)) / VmS


def df_AmL(T, Tg, alpha, VmL):
    return R_GAS * Tg * math.log(1.0 + alpha) * f_tau(T / Tg) / VmL


class Model:
    """The run constants of one GFA settings file ("constants" block) and the
    temperature-only quantities derived from them (GFA.cu:104-121)."""

    def __init__(self, k: dict):
        self.k = dict(k)
        T = k["T"]
        self.T = T
        self.M_phi = 22.1 * math.exp(-140.0e3 / (R_GAS * T))
        self.D_L = 2.0e-6 * math.exp(-1.11e5 / (R_GAS * T))
        self.D_S = 1.311e-6 * math.exp(-1.51e5 / (R_GAS * T))
        self.fS = f_S(T, k["V_m_S"])
        self.dfAm = df_AmL(T, k["T_g"], k["alpha"], k["V_m_L"])

    # pointwise parts of the six equations (GFA.cu:144-190)
    def mu(self, c, phi):
        return (1.0 - h(phi)) * dGL_dc(c, self.T) / self.k["V_m_L"]

    def mobility(self, phi, eta):
        hp, he = h(phi), h(eta)
        return ((1.0 - hp) * ((1.0 - he) * self.D_L + he * self.k["D_Am"]) + hp * self.D_S) / (R_GAS * self.T)

    def D(self, c, phi, eta):
        return self.mobility(phi, eta) * c * (1.0 - c)

    def dF_dphi(self, phi, eta, c):
        k = self.k
        return (h_prime(phi) * (self.fS - f_L(c, self.T, k["V_m_L"]) - h(eta) * self.dfAm)
                + k["w_phi"] * g_prime(phi) + 2.0 * k["w_ex"] * phi * eta * eta)

    def dF_deta(self, phi, eta):
        k = self.k
        return ((1.0 - h(phi)) * h_prime(eta) * self.dfAm + k["w_eta"] * g_prime(eta)
                + 2.0 * k["w_ex"] * phi * phi * eta)


def pad(f, bc):
    """One ghost layer per side; bc = (x_kind, y_kind), each "Periodic" or
    "NoFlux" (both faces of an axis alike, as BCFactory builds them)."""
    out = f
    for axis, kind in ((1, bc[0]), (0, bc[1])):
        width = [(0, 0), (0, 0)]
        width[axis] = (1, 1)
        out = np.pad(out, width, mode="wrap" if kind == "Periodic" else "edge")
    return out


def lap(f, dx, dy, bc):
    """(f[+1] - 2 f + f[-1]) / dx^2 + (same along y) / dy^2 with
    1/d^2 formed first (Equation.cu:40-80,141-143)."""
    p = pad(f, bc)
    idx2, idy2 = 1.0 / (dx * dx), 1.0 / (dy * dy)
    cx = p[1:-1, 1:-1]
    return (p[1:-1, 2:] - 2.0 * cx + p[1:-1, :-2]) * idx2 + (p[2:, 1:-1] - 2.0 * cx + p[:-2, 1:-1]) * idy2


def grad(f, axis, d, bc):
    """(f[+1] - f[-1]) * (0.5 / d) along axis 0 = x, 1 = y (Equation.cu:212-237)."""
    p = pad(f, bc)
    inv_2d = 0.5 / d
    if axis == 0:
        return (p[1:-1, 2:] - p[1:-1, :-2]) * inv_2d
    return (p[2:, 1:-1] - p[:-2, 1:-1]) * inv_2d


def grad_zero_ghosts(f, axis, d):
    """grad of a field whose ghost layer holds zeros (same central difference
    as grad)."""
    p = np.pad(f, 1, mode="constant", constant_values=0.0)
    inv_2d = 0.5 / d
    if axis == 0:
        return (p[1:-1, 2:] - p[1:-1, :-2]) * inv_2d
    return (p[2:, 1:-1] - p[:-2, 1:-1]) * inv_2d


def c_operator(D, mu, dx, dy, bc):
    """D lap(mu) + dD/dx dmu/dx + dD/dy dmu/dy (GFA.cu:160-165). mu's ghost
    layer is the BC fill (multi-step 2 refreshes it, Solver.cu:322). D's ghost
    layer is zero, whatever the BC: multi-step 3 fills D's ghosts and then
    recomputes D as a STEADY unknown, and computeRHS clears the whole stored
    array before the pointwise term writes the physical cells
    (Solver.cu:322-326, Equation.cu:265, TermPW.inl:28-36)."""
    return (D * lap(mu, dx, dy, bc) + grad_zero_ghosts(D, 0, dx) * grad(mu, 0, dx, bc)
            + grad_zero_ghosts(D, 1, dy) * grad(mu, 1, dy, bc))


def steps(m: Model, c, phi, eta, dx, dy, bc, dt, nsteps):
    """nsteps multi-step Euler steps in the reference's order: mu; D (twice:
    the second fills D's ghosts, then clears them again by recomputing D, see
    c_operator); c; phi reading the new c; eta reading the new phi. Returns
    new (c, phi, eta, mu, D)."""
    c, phi, eta = (np.array(a, dtype=np.float64, copy=True) for a in (c, phi, eta))
    mu = D = None
    k = m.k
    for _ in range(nsteps):
        mu = m.mu(c, phi)
        D = m.D(c, phi, eta)
        c = c + dt * c_operator(D, mu, dx, dy, bc)
        phi = phi + dt * (-m.M_phi * m.dF_dphi(phi, eta, c) + (m.M_phi * k["eps_sq"]) * lap(phi, dx, dy, bc))
        eta = eta + dt * (-k["M_eta"] * m.dF_deta(phi, eta) + (k["M_eta"] * k["beta_sq"]) * lap(eta, dx, dy, bc))
    return c, phi, eta, mu, D
