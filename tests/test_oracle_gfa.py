"""Pins of the Fe-B glass-formation oracle (oracle/gfa.py) by facts the
mathematics and the published model fix, not by restating its formulas:

* every SGTE Gibbs-energy piece meets its neighbour at the breakpoint in value
  (2e-7 relative: the SGTE decimals' own rounding leaves up to 1.3e-7) and
  slope (2e-4 J/mol/K), except the Fe pieces at 1811 K, whose
  jumps equal those of the reference's 773597/T against SGTE's 77359/T
  (GFA.cu:32,39);
* the Inden f(tau) decimals equal the closed form with p = 0.6, and f is
  continuous at tau = 1;
* dG_L/dc is the c-derivative of G_L; h' and g' are the derivatives of h and
  of g(x) = x^2 (1 - x)^2;
* mu, dF/dphi and dF/deta are the partial derivatives of one density
  f = h(phi) f_S + (1 - h(phi)) (f_L(c) + h(eta) df_Am) + w_phi g(phi)
  + w_eta g(eta) + w_ex phi^2 eta^2;
* the mobility reduces to D_L, D_S, D_Am over R T in the pure phases;
* lap and grad equal the pinned C oracle's stencils under both BCs, and the c
  operator converges at second order to div(D grad mu);
* the model header used by the gfa_2d program evaluates every piece to the
  oracle's values on the host (tests/cpp/gfa_thermo.cpp)."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import gfa as G
from oracle import oracle as O
from phix_amd import capi

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "phix_amd" / "_build" / "bin"

PIECES = {   # function, breakpoints
    "G_B_liquid": (G.G_B_liquid, (500.0, 2348.0)),
    "G_Fe_liquid": (G.G_Fe_liquid, (1811.0,)),
    "G_Fe_bcc": (G.G_Fe_bcc, (1811.0,)),
    "G_B_beta": (G.G_B_beta, (1100.0, 2348.0, 3000.0)),
}
# the piece below 1811 K carries (773597 - 77359)/T more than SGTE's: right minus left
FE_JUMP = -(773597.0 - 77359.0) / 1811.0
FE_SLOPE_JUMP = (773597.0 - 77359.0) / 1811.0 ** 2

# constants block of the reference's FeB_ex1 settings (tests/golden/settings_gfa_2d.jsonc)
FEB = {"M_eta": 1.0, "D_Am": 3.0e-12, "eps_sq": 2.86e-3, "beta_sq": 1.0e-4, "w_phi": 441.0, "w_eta": 10.0,
       "w_ex": 100.0, "T": 1000.0, "T_g": 710.0, "alpha": 0.45, "V_m_L": 2.64e-5, "V_m_S": 2.44e-5}


def _slope(f, T, side, d=1e-3):
    return (f(T) - f(T - d)) / d if side < 0 else (f(T + d) - f(T)) / d


@pytest.mark.parametrize("name", list(PIECES))
def test_sgte_pieces_meet_at_breakpoints(name):
    f, breaks = PIECES[name]
    for Tb in breaks:
        above = math.nextafter(Tb, math.inf)
        jump = f(above) - f(Tb)
        slope_jump = _slope(f, above, +1) - _slope(f, Tb, -1)
        expect, expect_slope = (FE_JUMP, FE_SLOPE_JUMP) if name.startswith("G_Fe") else (0.0, 0.0)
        # the left piece at Tb, the right piece just above it: the reference's
        # pieces switch on T <= Tb
        assert abs(jump - expect) <= 2e-7 * abs(f(Tb)), (name, Tb, jump)
        assert abs(slope_jump - expect_slope) <= 2e-4, (name, Tb, slope_jump)


def test_fe_pieces_jump_as_the_773597_coefficient_says():
    """With SGTE's 77359/T in place of 773597/T the Fe pieces are continuous:
    the jump is that one coefficient, not a wrong digit elsewhere."""
    for f in (G.G_Fe_liquid, G.G_Fe_bcc):
        fixed = lambda T, f=f: f(T) - (773597.0 - 77359.0) / T if T <= 1811.0 else f(T)
        assert abs(fixed(math.nextafter(1811.0, math.inf)) - fixed(1811.0)) <= 2e-7 * abs(f(1811.0))


def test_inden_decimals_are_the_closed_form():
    p = 0.6
    A = 518.0 / 1125.0 + 11692.0 / 15975.0 * (1.0 / p - 1.0)
    k = 474.0 / 497.0 * (1.0 / p - 1.0)

    def closed(tau):
        if tau <= 1.0:
            return 1.0 - (79.0 / (140.0 * p * tau) + k * (tau ** 3 / 6.0 + tau ** 9 / 135.0 + tau ** 15 / 600.0)) / A
        return -(tau ** -5 / 10.0 + tau ** -15 / 315.0 + tau ** -25 / 1500.0) / A

    for tau in (0.3, 0.6, 0.9, 1.0, 1.1, 1.408, 2.0, 3.5):
        assert abs(G.f_tau(tau) - closed(tau)) <= 1e-6 * max(1.0, abs(closed(tau))), tau
    assert abs(G.f_tau(math.nextafter(1.0, math.inf)) - G.f_tau(1.0)) <= 1e-6


@pytest.mark.parametrize("T", [450.0, 1000.0, 1900.0, 2500.0, 3100.0])
def test_dGL_dc_is_the_derivative_of_GL(T):
    c = np.array([0.05, 0.25, 0.33, 0.5, 0.8, 0.97])
    hc = 1e-6
    fd = (G.G_L(c + hc, T) - G.G_L(c - hc, T)) / (2 * hc)
    assert np.abs(fd - G.dGL_dc(c, T)).max() <= 1e-3
    # outside [1e-12, 1 - 1e-12] the clamp holds the value
    assert G.dGL_dc(np.array([-0.2]), T)[0] == G.dGL_dc(np.array([1e-12]), T)[0]
    assert G.G_L(np.array([1.5]), T)[0] == G.G_L(np.array([1.0 - 1e-12]), T)[0]


def test_interpolation_and_well_derivatives():
    x = np.linspace(-0.2, 1.2, 29)
    hx = 1e-6
    assert np.abs((G.h(x + hx) - G.h(x - hx)) / (2 * hx) - G.h_prime(x)).max() <= 1e-8
    g = lambda v: v * v * (1.0 - v) ** 2
    assert np.abs((g(x + hx) - g(x - hx)) / (2 * hx) - G.g_prime(x)).max() <= 1e-8
    assert G.h(0.0) == 0.0 and G.h(1.0) == 1.0 and G.h(0.5) == 0.5


def test_forces_are_derivatives_of_one_density():
    m = G.Model(FEB)
    k = FEB

    def g(v):
        return v * v * (1.0 - v) ** 2

    def density(c, phi, eta):
        return (G.h(phi) * m.fS + (1.0 - G.h(phi)) * (G.f_L(c, m.T, k["V_m_L"]) + G.h(eta) * m.dfAm)
                + k["w_phi"] * g(phi) + k["w_eta"] * g(eta) + k["w_ex"] * phi * phi * eta * eta)

    rng = np.random.default_rng(5)
    c = rng.uniform(0.1, 0.6, 50)
    phi = rng.uniform(-0.05, 1.05, 50)
    eta = rng.uniform(-0.05, 1.05, 50)
    e = 1e-6
    d_c = (density(c + e, phi, eta) - density(c - e, phi, eta)) / (2 * e)
    d_phi = (density(c, phi + e, eta) - density(c, phi - e, eta)) / (2 * e)
    d_eta = (density(c, phi, eta + e) - density(c, phi, eta - e)) / (2 * e)
    scale = np.abs(m.fS) + 1.0   # densities are ~1e9 J/m^3: finite differences carry ~1e-7 relative error
    assert np.abs(d_c - m.mu(c, phi)).max() <= 1e-7 * scale
    assert np.abs(d_phi - m.dF_dphi(phi, eta, c)).max() <= 1e-7 * scale
    assert np.abs(d_eta - m.dF_deta(phi, eta)).max() <= 1e-7 * scale
    assert np.abs(m.dF_dphi(phi, eta, c)).max() > 1e6   # not vacuous


def test_mobility_pure_phase_limits():
    m = G.Model(FEB)
    RT = G.R_GAS * FEB["T"]
    assert math.isclose(m.mobility(0.0, 0.0), m.D_L / RT, rel_tol=1e-15)
    assert math.isclose(m.mobility(1.0, 0.0), m.D_S / RT, rel_tol=1e-15)
    assert math.isclose(m.mobility(1.0, 1.0), m.D_S / RT, rel_tol=1e-15)
    assert math.isclose(m.mobility(0.0, 1.0), FEB["D_Am"] / RT, rel_tol=1e-15)
    # Arrhenius forms at 1000 K
    assert math.isclose(m.D_L, 2.0e-6 * math.exp(-1.11e5 / 8314.0), rel_tol=1e-14)


@pytest.mark.parametrize("bc", [("Periodic", "Periodic"), ("NoFlux", "Periodic"), ("NoFlux", "NoFlux")])
def test_stencils_equal_the_c_oracle(bc):
    nx, ny, dx, dy = 13, 9, 0.7, 1.3
    kind = {"Periodic": capi.PERIODIC, "NoFlux": capi.NOFLUX}
    g = capi.make_grid((nx, ny), d=(dx, dy), bc=[[kind[bc[0]]] * 2, [kind[bc[1]]] * 2])
    f = np.random.default_rng(3).standard_normal((ny, nx))
    s = np.zeros(capi.stored_shape(g))
    capi.physical(s, g)[0] = f
    O.bc_fill(g, s)
    ref = np.zeros_like(s)
    O.lap_accumulate(g, ref, s, 1.0)
    assert np.array_equal(capi.physical(ref, g)[0], G.lap(f, dx, dy, bc))
    for axis, d in ((0, dx), (1, dy)):
        ref = np.zeros_like(s)
        O.grad_accumulate(g, ref, s, axis, 1.0)
        assert np.array_equal(capi.physical(ref, g)[0], G.grad(f, axis, d, bc)), axis


def test_c_operator_converges_to_div_D_grad_mu():
    errs = []
    for n in (32, 64, 128):
        d = 2 * math.pi / n
        x = (np.arange(n) + 0.5) * d
        X, Y = np.meshgrid(x, x)           # (ny, nx)
        D = 1.0 + 0.5 * np.sin(X) * np.cos(Y)
        mu = np.cos(2 * X) + np.sin(Y)
        # div(D grad mu) = D lap mu + grad D . grad mu
        lap_mu = -4 * np.cos(2 * X) - np.sin(Y)
        exact = D * lap_mu + (0.5 * np.cos(X) * np.cos(Y)) * (-2 * np.sin(2 * X)) + \
            (-0.5 * np.sin(X) * np.sin(Y)) * np.cos(Y)
        # interior cells: D's zero ghost layer enters only the face cells
        errs.append(np.abs(G.c_operator(D, mu, d, d, ("Periodic", "Periodic")) - exact)[1:-1, 1:-1].max())
    assert errs[0] / errs[1] > 3.8 and errs[1] / errs[2] > 3.9, errs


@pytest.mark.parametrize("bc", ["NoFlux", "Periodic"])
def test_c_operator_reads_zero_D_beyond_the_faces(bc):
    """The reference's multi-step 3 leaves D's ghost layer at zero for every
    BC, so at a face cell grad D reads D = 0 outside while grad mu and lap mu
    read mu's BC fill. Known answer for D = 1, mu = x along x (dx = 1, 5 x 4
    cells): interior 0; cell 0: lap mu = (mu1 - 2 mu0 + ghost)/dx^2 and
    dD/dx dmu/dx = (1 - 0)/2 * (mu1 - ghost)/2."""
    nx, ny = 5, 4
    mu = np.tile(np.arange(nx, dtype=float), (ny, 1))
    D = np.ones((ny, nx))
    got = G.c_operator(D, mu, 1.0, 1.0, (bc, bc))
    g0 = mu[0, 0] if bc == "NoFlux" else mu[0, -1]       # mu's ghost at x = -1
    g1 = mu[0, -1] if bc == "NoFlux" else mu[0, 0]       # mu's ghost at x = nx
    lo = (mu[0, 1] - 2 * mu[0, 0] + g0) + 0.5 * (mu[0, 1] - g0) / 2
    hi = (g1 - 2 * mu[0, -1] + mu[0, -2]) + (-0.5) * (g1 - mu[0, -2]) / 2
    assert np.array_equal(got[:, 1:-1], np.zeros((ny, nx - 2)))
    assert np.allclose(got[:, 0], lo, rtol=0, atol=1e-15) and np.allclose(got[:, -1], hi, rtol=0, atol=1e-15)
    # y faces: grad D along y reads the zero ghosts, but mu is constant in y
    assert np.array_equal(got[0, 1:-1], np.zeros(nx - 2))


def test_model_header_matches_oracle_on_host():
    """gfa_thermo prints the header's functions (host evaluation); each equals
    the oracle within a few ulps of the terms' magnitude."""
    r = subprocess.run([str(BIN / "gfa_thermo")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    fns = {"G_B_liquid": G.G_B_liquid, "G_Fe_liquid": G.G_Fe_liquid, "L_BFe_liquid": G.L_BFe_liquid,
           "G_Fe_bcc": G.G_Fe_bcc, "G_B_beta": G.G_B_beta, "deltaG_Fe3B": G.deltaG_Fe3B, "f_tau": G.f_tau,
           "h": G.h, "h_prime": G.h_prime, "g_prime": G.g_prime,
           "dGL_dc": lambda c, T: float(G.dGL_dc(np.array([c]), T)[0]),
           "G_L": lambda c, T: float(G.G_L(np.array([c]), T)[0]),
           "f_S": G.f_S, "df_AmL": G.df_AmL}
    seen = set()
    for line in r.stdout.splitlines():
        name, *vals = line.split()
        args, got = [float(v) for v in vals[:-1]], float(vals[-1])
        want = fns[name](*args)
        scale = max(abs(want), 1.0) if name in ("h", "h_prime", "g_prime", "f_tau") else max(abs(want), 1e4)
        assert abs(got - want) <= 1e-13 * scale, (name, args, got, want)
        seen.add(name)
    assert seen == set(fns)


def test_fused_entry_point_rejects_bad_arguments():
    """phx_gfa_euler_steps checks its arguments before any device work:
    3D grids, FIXED faces, ghost layers beyond the axis, null pointers, a
    non-finite constant, negative nsteps and a missing workspace are
    PHX_ERR_INVALID; nsteps = 0 returns PHX_OK without reading the fields."""
    import ctypes as C
    lib = capi.lib()
    p = capi.GfaParams(dt=1e-7, **FEB)
    g = capi.make_grid((8, 6), d=(1e-5, 1e-5), bc=capi.NOFLUX)
    x = C.c_void_p(16)   # never dereferenced: every call below fails or returns first
    inv = capi.PHX_ERR_INVALID

    def call(grid, *ptrs, params=p, nsteps=1):
        return lib.phx_gfa_euler_steps(C.byref(grid), *ptrs, C.byref(params), nsteps, None)

    assert lib.phx_gfa_workspace_bytes(C.byref(g)) == 3 * 8 * (10 * 8 * 3)
    assert call(capi.make_grid((8, 6, 4), bc=capi.NOFLUX), x, x, x, x) == inv
    assert call(capi.make_grid((8, 6), bc=[[capi.FIXED] * 2, [capi.NOFLUX] * 2]), x, x, x, x) == inv
    assert call(capi.make_grid((8, 2), ghost=2, bc=[[capi.NOFLUX] * 2, [capi.PERIODIC] * 2]), x, x, x, x) != capi.PHX_OK
    assert call(g, None, x, x, x) == inv
    assert call(g, x, x, x, None) == inv
    assert call(g, x, x, x, x, nsteps=-1) == inv
    bad = capi.GfaParams(dt=1e-7, **dict(FEB, T=float("nan")))
    assert call(g, x, x, x, x, params=bad) == inv
    assert "non-finite" in capi.lib().phx_last_error().decode()
    assert call(g, x, x, x, None, nsteps=0) == capi.PHX_OK
