"""Pins of oracle/dsl.py (the numpy restatement of the composite DSL forms and
of the CH+AC program) by what the mathematics fixes, not by retyping it:

* gx, gy, lap: closed forms on Fourier modes (periodic) and on the cosine
  modes cos(pi m (i + 1/2) / n), whose mirror symmetry about the faces makes
  them exact eigenfunctions of the noflux stencils.
* composite_forms: every form on inputs whose result has a closed form.
* chac_steps: its pointwise forces are the partial derivatives of one free
  energy density (checked by finite differences), a uniform state follows an
  independent scalar Euler integration of that density's gradient flow, the
  linearised c equation amplifies each Fourier mode exactly by
  1 - dt M lam (2 rho^2 + kappa_c lam), and sum c is conserved."""
import math

import numpy as np
import pytest

from oracle import dsl

NX, NY, DX, DY = 24, 20, 0.5, 0.25


def mode(kind, mx, my, phase=0.0):
    """A mode on the (NY, NX) cell-centre grid: periodic cos(2 pi (mx i / NX +
    my j / NY) + phase), or the noflux cosine product cos(pi mx (i + 1/2) / NX)
    cos(pi my (j + 1/2) / NY). Returns (field, theta_x step, theta_y step)."""
    i = np.arange(NX)[None, :]
    j = np.arange(NY)[:, None]
    if kind == "periodic":
        hx, hy = 2 * math.pi * mx / NX, 2 * math.pi * my / NY
        return np.cos(hx * i + hy * j + phase) + 0 * i * j, hx, hy
    hx, hy = math.pi * mx / NX, math.pi * my / NY
    return np.cos(hx * (i + 0.5)) * np.cos(hy * (j + 0.5)), hx, hy


@pytest.mark.parametrize("mx,my", [(1, 0), (3, 2), (5, 7), (11, 9)])
def test_periodic_stencils_on_fourier_modes(mx, my):
    f, hx, hy = mode("periodic", mx, my, 0.3)
    i = np.arange(NX)[None, :]
    j = np.arange(NY)[:, None]
    s = np.sin(hx * i + hy * j + 0.3)
    assert np.allclose(dsl.gx(f, DX, "periodic"), -s * math.sin(hx) / DX, rtol=0, atol=1e-13)
    assert np.allclose(dsl.gy(f, DY, "periodic"), -s * math.sin(hy) / DY, rtol=0, atol=1e-13)
    lam = 4 * math.sin(hx / 2) ** 2 / DX ** 2 + 4 * math.sin(hy / 2) ** 2 / DY ** 2
    assert np.allclose(dsl.lap(f, DX, DY, "periodic"), -lam * f, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mx,my", [(0, 0), (1, 0), (0, 3), (4, 5), (23, 19)])
def test_noflux_lap_on_cosine_modes(mx, my):
    f, hx, hy = mode("noflux", mx, my)
    lam = 4 * math.sin(hx / 2) ** 2 / DX ** 2 + 4 * math.sin(hy / 2) ** 2 / DY ** 2
    assert np.allclose(dsl.lap(f, DX, DY, "noflux"), -lam * f, rtol=0, atol=1e-12)
    # the face cells' ghosts are the face values: gx at the faces is the
    # central difference of the mirrored mode as well
    i = np.arange(NX)[None, :]
    j = np.arange(NY)[:, None]
    want = -np.sin(hx * (i + 0.5)) * math.sin(hx) / DX * np.cos(hy * (j + 0.5))
    assert np.allclose(dsl.gx(f, DX, "noflux"), want, rtol=0, atol=1e-13)


def test_pad_ghosts():
    f = np.arange(NX * NY, dtype=float).reshape(NY, NX)
    p = dsl.pad(f, "periodic")
    assert p[0, 0] == f[-1, -1] and p[-1, -1] == f[0, 0] and p[5, 0] == f[4, -1] and p[0, 7] == f[-1, 6]
    q = dsl.pad(f, "noflux")
    assert q[0, 0] == f[0, 0] and q[-1, -1] == f[-1, -1] and q[5, 0] == f[4, 0] and q[-1, 7] == f[-1, 6]
    with pytest.raises(ValueError):
        dsl.pad(f, "fixed")


@pytest.mark.parametrize("kind", ["periodic", "noflux"])
def test_composite_forms_closed_forms(kind):
    """Each form on inputs with a closed-form result."""
    one = np.ones((NY, NX))
    zero = np.zeros((NY, NX))
    m1, hx, hy = mode(kind, 3, 2)
    lam1 = 4 * math.sin(hx / 2) ** 2 / DX ** 2 + 4 * math.sin(hy / 2) ** 2 / DY ** 2
    # a = mode, b = 1: lap(a b) = -lam a; (a b) lap(b) = 0; a*b enters as a
    r = dsl.composite_forms(m1, one, zero, zero, DX, DY, kind)
    assert np.allclose(r["lap_ab"], -lam1 * m1, rtol=0, atol=1e-11)
    assert np.array_equal(r["ab_lapb"], zero)
    assert np.allclose(r["mixed"], 2.0 * dsl.gx(m1, DX, kind) - m1, rtol=0, atol=1e-13)
    # a = 1, b = mode: (a b) lap(b) = -lam b^2
    r = dsl.composite_forms(one, m1, zero, zero, DX, DY, kind)
    assert np.allclose(r["ab_lapb"], -lam1 * m1 * m1, rtol=0, atol=1e-11)
    if kind == "periodic":
        i = np.arange(NX)[None, :]
        j = np.arange(NY)[:, None]
        th = hx * i + hy * j
        # a = cos th: a^2 = (1 + cos 2th)/2, so gx(a^2) = -sin(2th) sin(2hx) / (2 dx)
        a = np.cos(th)
        r = dsl.composite_forms(a, one, zero, zero, DX, DY, kind)
        assert np.allclose(r["gradx_sq"], -np.sin(2 * th) * math.sin(2 * hx) / (2 * DX), rtol=0, atol=1e-12)
        # a b = cos th: gy = -sin th sin(hy)/dy; div grad = -(sin^2 hx / dx^2 + sin^2 hy / dy^2) cos th
        assert np.allclose(r["grady_ab"], -np.sin(th) * math.sin(hy) / DY, rtol=0, atol=1e-13)
        wide = (math.sin(hx) / DX) ** 2 + (math.sin(hy) / DY) ** 2
        assert np.allclose(r["divgrad_ab"], -wide * a, rtol=0, atol=1e-11)
        # mixed with a = 0: div(v) of v = (cos th, sin th)
        r = dsl.composite_forms(zero, zero, np.cos(th), np.sin(th), DX, DY, kind)
        want = -np.sin(th) * math.sin(hx) / DX + np.cos(th) * math.sin(hy) / DY
        assert np.allclose(r["mixed"], want, rtol=0, atol=1e-13)
    else:
        # a b = noflux cosine mode: gx(gx) at the faces reads gx's own ghost
        # (the face value), so only the interior has the wide-stencil form
        r = dsl.composite_forms(m1, one, zero, zero, DX, DY, kind)
        wide = (math.sin(hx) / DX) ** 2 + (math.sin(hy) / DY) ** 2
        assert np.allclose(r["divgrad_ab"][1:-1, 1:-1], -wide * m1[1:-1, 1:-1], rtol=0, atol=1e-11)


def test_composite_inputs_are_the_driver_fields():
    a, b, v0, v1 = dsl.composite_inputs()
    assert a.shape == (NY, NX)
    x0, y0 = 0.25, 0.125          # first cell centre: dx / 2, dy / 2
    assert a[0, 0] == 1.5 + math.sin(x0) * math.cos(2 * y0)
    assert b[0, 1] == math.cos(0.7 * 0.75 - 1.3 * y0)
    assert v0[2, 0] == x0 * math.sin(0.625) and v1[0, 3] == math.exp(-0.1 * 1.75) * y0


# ------------------------------------------------------------------ CH + AC
K = {"M": 5.0, "rho": math.sqrt(2.0), "ca": 0.3, "cb": 0.7, "kappa_c": 3.0, "kappa_eta": 3.0, "L": 5.0, "w": 1.0}


def density(c, eta, k):
    """Bulk free energy density the CH+AC forces descend, written from the
    model's parts: the two parabolic phases weighted by the interpolation
    h = eta^3 (10 - 15 eta + 6 eta^2), and the double well w eta^2 (1 - eta)^2."""
    h = eta ** 3 * (10.0 - 15.0 * eta + 6.0 * eta * eta)
    return k["rho"] ** 2 * ((1.0 - h) * (c - k["ca"]) ** 2 + h * (c - k["cb"]) ** 2) + k["w"] * eta ** 2 * (1 - eta) ** 2


def test_chac_forces_are_derivatives_of_one_density():
    """One step on uniform fields: c stays put (lap of a uniform mu is 0) and
    eta moves by -dt L df/deta; the step's mu equals df/dc (both by central
    finite differences of `density`)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        c0, e0 = rng.uniform(0.1, 0.9), rng.uniform(-0.2, 1.2)
        c = np.full((6, 7), c0)
        eta = np.full((6, 7), e0)
        dt = 1e-3
        c1, e1, mu = dsl.chac_steps(c, eta, K, dt, 1)
        hd = 1e-6
        dfdc = (density(c0 + hd, e0, K) - density(c0 - hd, e0, K)) / (2 * hd)
        dfde = (density(c0, e0 + hd, K) - density(c0, e0 - hd, K)) / (2 * hd)
        assert np.array_equal(c1, c)
        assert np.allclose(mu, dfdc, rtol=1e-8, atol=1e-8)
        assert np.allclose((e1 - eta) / dt, -K["L"] * dfde, rtol=1e-7, atol=1e-7)


def test_chac_uniform_state_follows_an_independent_scalar_integration():
    """A uniform state for 400 steps against a scalar Euler integration of
    d eta / dt = -L df/deta with the derivative taken numerically."""
    c0, e0, dt = 0.62, 0.35, 1e-3
    c1, e1, _ = dsl.chac_steps(np.full((4, 5), c0), np.full((4, 5), e0), K, dt, 400)
    e = e0
    for _ in range(400):
        hd = 1e-7
        e = e - dt * K["L"] * (density(c0, e + hd, K) - density(c0, e - hd, K)) / (2 * hd)
    assert np.allclose(c1, c0, rtol=0, atol=0)
    assert np.allclose(e1, e, rtol=0, atol=1e-7)
    assert abs(e1[0, 0] - e0) > 0.1          # it moved (towards the phase closer to c)


@pytest.mark.parametrize("m", [(1, 0), (2, 3), (5, 1)])
def test_chac_linearised_c_modes_amplify_exactly(m):
    """eta = 0 makes h = 0 and mu = 2 rho^2 (c - ca) - kappa_c lap(c): linear,
    so a Fourier mode of c is multiplied per step by 1 - dt M lam (2 rho^2 +
    kappa_c lam), lam = 4 sin^2(pi m_x / n_x) + 4 sin^2(pi m_y / n_y), while
    eta stays 0 (every force on it carries a factor eta)."""
    nx, ny, dt, steps = 16, 12, 1e-3, 25
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    eps = 1e-3
    f = np.cos(2 * math.pi * (m[0] * i / nx + m[1] * j / ny))
    c = K["ca"] + eps * f
    c1, e1, _ = dsl.chac_steps(c, np.zeros((ny, nx)), K, dt, steps)
    lam = 4 * math.sin(math.pi * m[0] / nx) ** 2 + 4 * math.sin(math.pi * m[1] / ny) ** 2
    amp = (1 - dt * K["M"] * lam * (2 * K["rho"] ** 2 + K["kappa_c"] * lam)) ** steps
    assert np.allclose(c1, K["ca"] + eps * amp * f, rtol=0, atol=1e-14)
    assert np.array_equal(e1, np.zeros((ny, nx)))


@pytest.mark.parametrize("kind", ["periodic", "noflux"])
def test_chac_conserves_mass(kind):
    rng = np.random.default_rng(9)
    c = 0.5 + rng.uniform(-0.05, 0.05, (18, 22))
    eta = rng.uniform(0.0, 1.0, (18, 22))
    c1, e1, _ = dsl.chac_steps(c, eta, K, 1e-3, 200, mode=kind)
    assert abs(c1.sum() - c.sum()) <= 1e-12 * c.size
    assert np.abs(e1 - eta).max() > 1e-2
