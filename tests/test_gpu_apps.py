"""GPU tests of the config-driven solver programs (phix_amd/apps): ch_2d and
ch_ac_2d run from a settings file in a scratch working directory the way the
reference's CH_2D / CH_AC_2D programs do (settings/settings.jsonc,
settings/initial_field/*.field, output/<name>_<step>.*).

* ch_2d: the fused Euler plan against the FP64 oracle (<= 1e-12), periodic
  and no-flux; restart from a written step reproduces the uninterrupted run
  bit for bit; the DAT output holds the same values as the BINARY one; the
  optional FP32 bulk setting stays within FP32 accuracy.
* ch_ac_2d: the coupled CH + AC system on the generic device path against a
  numpy FP64 restatement of the reference program's equations and multi-step
  order (CH+AC_double-well.cu:95-136, Solver.cu multi-step: mu, then c, then
  eta reading the updated c); tolerance 1e-11 (different summation order and
  contraction of the per-term kernels)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, write_field, read_field

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "phix_amd" / "_build" / "bin"

CH_CONST = {"M": 5.0, "kappa": 2.0, "rho": 5.0, "ca": 0.3, "cb": 0.7}   # settings_ch_2d.jsonc
CHAC_CONST = {"M": 5.0, "rho": 1.4142135623730951, "ca": 0.3, "cb": 0.7, "kappa_c": 3.0, "kappa_eta": 3.0,
              "L": 5.0, "w": 1.0}                                         # settings_ch_ac_2d.jsonc


def settings(nx, ny, const, nsteps, start="initial_field", dt=0.001, bcs=("Periodic",) * 4, fmt="ALL",
             print_every=100, write_every=100, extra=None):
    s = {"mesh": {"nx": nx, "ny": ny, "dx": 1.0, "dy": 1.0, "x0": 0.0, "y0": 0.0},
         "initialize": {"start_from": start, "dt": dt, "nSteps": nsteps},
         "boundary_conditions": dict(zip(("x_min", "x_max", "y_min", "y_max"), bcs)),
         "constants": const,
         "output": {"print_interval": print_every, "write_interval": write_every, "format": fmt}}
    if extra:
        s["initialize"].update(extra)
    return "// generated for the test\n" + json.dumps(s, indent=2) + "\n"


def prepare(work, text, fields):
    (work / "settings" / "initial_field").mkdir(parents=True, exist_ok=True)
    (work / "settings" / "settings.jsonc").write_text(text)
    for name, phys in fields.items():
        write_field(work / "settings" / "initial_field" / f"{name}.field", phys, name=name)


def run(exe, work, *args, timeout=300):
    print(f"running {exe} in {work}", flush=True)
    r = subprocess.run([str(BIN / exe), *map(str, args)], capture_output=True, text=True, timeout=timeout, cwd=work)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def dat_values(path):
    rows = np.loadtxt(path, comments="#")
    return rows[:, 3]


@pytest.mark.parametrize("bcs", [("Periodic",) * 4, ("NoFlux", "NoFlux", "Periodic", "Periodic")])
def test_ch_2d_matches_oracle_and_restarts(dev, tmp_path, bcs):
    nx, ny, nsteps = 64, 48, 300
    c0 = init.ch_field((nx, ny), 11, 0.5, 0.02)
    work = tmp_path / "run"
    prepare(work, settings(nx, ny, CH_CONST, nsteps, bcs=bcs, write_every=150), {"c": c0, "mu": np.zeros_like(c0)})
    out = run("ch_2d", work)
    assert "solver plan: fused CH Euler" in out
    assert "[progress] step=300" in out and "Done." in out
    for step in (0, 150, 300):
        for ext in ("field", "dat", "vts"):
            assert (work / "output" / f"c_{step}.{ext}").exists(), (step, ext)

    P, N = capi.PERIODIC, capi.NOFLUX
    bc = [[P, P], [P, P]] if bcs[0] == "Periodic" else [[N, N], [P, P]]
    g = capi.make_grid((nx, ny), bc=bc)
    p = ch_params(rho=5.0, ca=0.3, cb=0.7, kappa=2.0, M=5.0, dt=0.001)
    c = init.to_stored(c0, g)
    mu = np.zeros_like(c)
    O.ch_euler_steps(g, p, c, mu, 150)
    _, c150 = read_field(work / "output" / "c_150.field")
    assert np.abs(c150[0] - capi.physical(c, g)).max() <= 1e-12
    O.ch_euler_steps(g, p, c, mu, 150)
    _, c300 = read_field(work / "output" / "c_300.field")
    err = np.abs(c300[0] - capi.physical(c, g)).max()
    print(f"ch_2d {bcs[0]}: max |c - oracle| after {nsteps} steps = {err:.3e}")
    assert err <= 1e-12
    # DAT carries the same values (17 significant digits)
    assert np.array_equal(dat_values(work / "output" / "c_300.dat"), c300[0].ravel())

    # warm start: "160" resolves to the nearest written step, 150
    first = (work / "output" / "c_300.field").read_bytes()
    (work / "settings" / "settings.jsonc").write_text(
        settings(nx, ny, CH_CONST, nsteps, start="160", bcs=bcs, write_every=150))
    out = run("ch_2d", work)
    assert "warm start: loading output/c_150.field" in out
    assert "Resuming Cahn-Hilliard simulation from step 150" in out
    assert (work / "output" / "c_300.field").read_bytes() == first


def test_ch_2d_fp32_bulk(dev, tmp_path):
    nx, ny, nsteps = 64, 64, 200
    c0 = init.ch_field((nx, ny), 12, 0.5, 0.02)
    work = tmp_path / "run"
    prepare(work, settings(nx, ny, CH_CONST, nsteps, fmt="BINARY", write_every=200,
                           extra={"bulk_precision": "FP32"}), {"c": c0, "mu": np.zeros_like(c0)})
    out = run("ch_2d", work, "settings/settings.jsonc")
    assert "solver plan: fused CH Euler" in out
    g = capi.make_grid((nx, ny), bc=capi.PERIODIC)
    c = init.to_stored(c0, g)
    O.ch_euler_steps(g, ch_params(rho=5.0, ca=0.3, cb=0.7, kappa=2.0, M=5.0, dt=0.001), c, np.zeros_like(c), nsteps)
    _, cc = read_field(work / "output" / f"c_{nsteps}.field")
    err = np.abs(cc[0] - capi.physical(c, g)).max()
    print(f"ch_2d FP32 bulk: max |c - oracle| = {err:.3e}")
    assert 0 < err <= 1e-5   # FP32 arithmetic, so not bit-equal to FP64


def test_ch_ac_2d_matches_oracle(dev, tmp_path):
    nx, ny, nsteps = 48, 40, 200
    c0 = init.ch_field((nx, ny), 13, 0.5, 0.05)
    eta0 = init.ch_field((nx, ny), 14, 0.5, 0.3)
    work = tmp_path / "run"
    prepare(work, settings(nx, ny, CHAC_CONST, nsteps, fmt="BINARY", print_every=50, write_every=200),
            {"c": c0, "eta": eta0, "mu": np.zeros_like(c0)})
    out = run("ch_ac_2d", work)
    assert "solver plan: generic" in out and "Done." in out
    from oracle import dsl
    c_ref, eta_ref, _ = dsl.chac_steps(c0[0], eta0[0], CHAC_CONST, 0.001, nsteps)
    _, cc = read_field(work / "output" / f"c_{nsteps}.field")
    _, ee = read_field(work / "output" / f"eta_{nsteps}.field")
    ec, ee_ = np.abs(cc[0, 0] - c_ref).max(), np.abs(ee[0, 0] - eta_ref).max()
    print(f"ch_ac_2d: max |c - oracle| = {ec:.3e}, max |eta - oracle| = {ee_:.3e}")
    assert np.abs(eta_ref - eta0[0]).max() > 1e-3   # the run moved the fields
    assert ec <= 1e-11 and ee_ <= 1e-11


def _feb_constants():
    """The constants block of the reference's FeB_ex1 settings (JSONC: the
    // comments are stripped before json parses it)."""
    import re
    text = (ROOT / "tests" / "golden" / "settings_gfa_2d.jsonc").read_text(encoding="utf-8")
    return json.loads(re.sub(r"//[^\n]*", "", text))["constants"]


@pytest.mark.parametrize("state", ["nucleus", "noisy"])
@pytest.mark.parametrize("plan", [None, "generic"])
@pytest.mark.parametrize("bc", ["NoFlux", "Periodic"])
def test_gfa_2d_matches_oracle_and_restarts(dev, tmp_path, bc, plan, state):
    """gfa_2d (GFA.cu) on its default fused step and (plan "generic") on the
    per-term path against the FP64 numpy oracle/gfa.py, the reference's FeB_ex1
    constants, 48 x 40 cells of 20 um, dt = 0.2 us (inside the explicit limit
    of the eta gradient term), 200 steps. States: "nucleus", a B-rich crystal
    nucleus (phi = 1, eta = 0, c = 0.33 + U(-0.01, 0.01), radius 8 cells) in
    an amorphous matrix (phi = 0, eta = 1, c = 0.25); "noisy", c, phi and eta
    drawn with noise on every cell up to the faces, so mu and D vary at the
    faces and D's zero ghost layer (the reference's multi-step 3 clears it)
    enters the c update there. Tolerance 1e-10 max-abs on c, phi, eta
    (order-1 fields; different summation order, FMA contraction and the
    device log). A warm start from step 100 reproduces step 200 bit for bit."""
    from oracle import gfa as G
    k = _feb_constants()
    nx, ny, dx, dt, nsteps = 48, 40, 2e-5, 2e-7, 200
    if state == "nucleus":
        yy, xx = np.mgrid[0:ny, 0:nx]
        inside = (xx + 0.5 - nx / 2) ** 2 + (yy + 0.5 - ny / 2) ** 2 <= 8.0 ** 2
        noise = np.random.default_rng(21).uniform(-0.01, 0.01, (ny, nx))
        c0 = np.where(inside, 0.33 + noise, 0.25)
        phi0 = np.where(inside, 1.0, 0.0)
        eta0 = np.where(inside, 0.0, 1.0)
    else:
        rng = np.random.default_rng(22)
        c0 = 0.27 + rng.uniform(-0.03, 0.03, (ny, nx))
        phi0 = rng.uniform(0.0, 0.01, (ny, nx))
        eta0 = 1.0 - rng.uniform(0.0, 0.01, (ny, nx))
    text = settings(nx, ny, k, nsteps, dt=dt, bcs=(bc,) * 4, fmt="BINARY", print_every=50, write_every=100,
                    extra={"solver": plan} if plan else None)
    text = text.replace('"dx": 1.0', f'"dx": {dx!r}').replace('"dy": 1.0', f'"dy": {dx!r}')
    work = tmp_path / "run"
    prepare(work, text, {"c": c0[None], "phi": phi0[None], "eta": eta0[None], "mu": np.zeros((1, ny, nx))})
    out = run("gfa_2d", work)
    assert "Starting GFA simulation (200 steps" in out and "Done." in out
    assert ("solver plan: generic multi-step" if plan else "solver plan: fused GFA step") in out
    assert "[progress] step=200" in out

    m = G.Model(k)
    c, phi, eta, _, _ = G.steps(m, c0, phi0, eta0, dx, dx, (bc, bc), dt, nsteps)
    errs = {}
    for name, ref in (("c", c), ("phi", phi), ("eta", eta)):
        _, got = read_field(work / "output" / f"{name}_{nsteps}.field")
        errs[name] = np.abs(got[0, 0] - ref).max()
    print(f"gfa_2d {bc} {plan or 'fused'}: max |GPU - oracle| after {nsteps} steps: " +
          ", ".join(f"{n} {e:.3e}" for n, e in errs.items()))
    assert max(errs.values()) <= 1e-10, errs
    # the run moved every field
    assert np.abs(c - c0).max() > 1e-3 and np.abs(phi - phi0).max() > 1e-4
    assert np.abs(eta - eta0).max() > (0.5 if state == "nucleus" else 1e-3)
    if state == "noisy":
        # D's ghost layer matters here: the same steps with D's ghosts BC-filled end elsewhere
        cf = c0.copy()
        for _ in range(5):
            mu_, D_ = m.mu(cf, phi0), m.D(cf, phi0, eta0)
            cf = cf + dt * (D_ * G.lap(mu_, dx, dx, (bc, bc)) + G.grad(D_, 0, dx, (bc, bc)) * G.grad(mu_, 0, dx, (bc, bc))
                            + G.grad(D_, 1, dx, (bc, bc)) * G.grad(mu_, 1, dx, (bc, bc)))
        c5 = G.steps(m, c0, phi0, eta0, dx, dx, (bc, bc), dt, 5)[0]
        assert np.abs(cf - c5).max() > 1e-9

    finals = {n: (work / "output" / f"{n}_{nsteps}.field").read_bytes() for n in ("c", "phi", "eta")}
    (work / "settings" / "settings.jsonc").write_text(text.replace('"start_from": "initial_field"',
                                                                   '"start_from": "100"'))
    out = run("gfa_2d", work)
    assert "Resuming GFA simulation from step 100" in out
    for n, b in finals.items():
        assert (work / "output" / f"{n}_{nsteps}.field").read_bytes() == b, n
