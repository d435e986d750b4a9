"""GPU tests of the drop-in C++ API (include/phix/*): programs written the way
the reference's applications are (tests/cpp/*.cpp) run the fused plan, the
generic per-term plan, the host path and FP32 bulk; their .field outputs are
compared with the FP64 oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, ac_params, write_field, write_vfield, read_field

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "phix_amd" / "_build" / "bin"


def run(args, timeout=300):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_fieldops_known_answers_gpu(dev):
    out = run([BIN / "fieldops", "gpu"])
    assert "FAIL" not in out


def test_ch_app_euler(dev, tmp_path):
    """The CH double-well app (config 1 sizes, 200 steps) through Solver's
    multi-step constructor: fused == generic == host == oracle (<= 1e-12)."""
    n = (128, 128)
    phys = init.ch_field(n, 1, 0.5, 0.01)
    write_field(tmp_path / "in.field", phys)
    out = run([BIN / "ch_solver", "euler", tmp_path / "in.field", tmp_path, 128, 128, 200])
    assert "plan fused: fused CH Euler" in out
    assert "plan generic: generic" in out
    assert "prev==curr fused: 1" in out
    g = capi.make_grid(n, bc=capi.PERIODIC)
    p = ch_params(kappa=0.5, dt=0.01)
    c = init.to_stored(phys, g)
    mu = np.zeros_like(c)
    O.ch_euler_steps(g, p, c, mu, 200)
    for tag in ("fused", "generic", "cpu"):
        _, cc = read_field(tmp_path / f"euler_{tag}_c.field")
        _, mm = read_field(tmp_path / f"euler_{tag}_mu.field")
        assert np.abs(cc[0] - capi.physical(c, g)).max() <= 1e-12, tag
        assert np.abs(mm[0] - capi.physical(mu, g)).max() <= 1e-12, tag


@pytest.mark.parametrize("shape,bcmode", [((40, 32, 28), "noflux"), ((48, 40, 1), "periodic"), ((36, 30, 26), "mixed")])
def test_ch_rk4_solver(dev, tmp_path, shape, bcmode):
    nx, ny, nz = shape
    n = (nx, ny) if nz == 1 else (nx, ny, nz)
    phys = init.ch_field(n, 2, 0.4, 0.02)
    write_field(tmp_path / "in.field", phys)
    steps = 40
    out = run([BIN / "ch_solver", "rk4", tmp_path / "in.field", tmp_path, nx, ny, nz, steps, bcmode])
    assert "plan fused: fused CH RK4" in out and "plan fp32: fused CH RK4" in out
    assert "plan generic: generic" in out
    assert "prev==curr fused: 1 callbacks 5" in out
    P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED
    dim = len(n)
    if bcmode == "noflux":
        bc, vals = N, 0.0
    elif bcmode == "periodic":
        bc, vals = P, 0.0
    else:
        bc = [[F, N], [P, P], [N, N]][:dim]
        vals = [[0.45, 0], [0, 0], [0, 0]][:dim]
    g = capi.make_grid(n, bc=bc, bc_value=vals)
    c = init.to_stored(phys, g)
    O.ch_rk4_steps(g, ch_params(kappa=1.0, dt=0.005), c, steps)
    ref = capi.physical(c, g)
    for tag in ("fused", "generic", "cpu"):
        _, cc = read_field(tmp_path / f"rk4_{tag}_c.field")
        err = np.abs(cc[0] - ref).max()
        assert err <= 1e-12, (tag, err)
    _, c32 = read_field(tmp_path / "rk4_fp32_c.field")
    err32 = np.abs(c32[0] - ref).max()
    print(f"fp32 bulk via Solver: max-abs {err32:.3e}")
    assert err32 <= 2e-6


@pytest.mark.parametrize("nf,bcmode", [(8, "periodic"), (3, "noflux")])
def test_ac_vector_solver(dev, tmp_path, nf, bcmode):
    n = (64, 48)
    a = init.ac_fields(n, nf, 3)
    write_vfield(tmp_path / "in.vfield", a)
    steps = 30
    out = run([BIN / "ac_solver", tmp_path / "in.vfield", tmp_path, 64, 48, steps, bcmode])
    assert "plan fused: fused AC RK4" in out and "plan generic: generic" in out
    g = capi.make_grid(n, bc=capi.PERIODIC if bcmode == "periodic" else capi.NOFLUX)
    etas = [init.to_stored(a[f], g) for f in range(nf)]
    O.ac_rk4_steps(g, ac_params(), etas, steps)
    for tag in ("fused", "generic", "cpu"):
        _, got = read_field(tmp_path / f"ac_{tag}.vfield")
        err = max(np.abs(got[f] - capi.physical(etas[f], g)).max() for f in range(nf))
        assert err <= 1e-12, (tag, err)
    _, got = read_field(tmp_path / "ac_fp32.vfield")
    err = max(np.abs(got[f] - capi.physical(etas[f], g)).max() for f in range(nf))
    assert err <= 1e-5


@pytest.mark.parametrize("mode", ["periodic", "noflux"])
def test_composite_forms_device_path(dev, tmp_path, mode):
    """Composite DSL forms on the device path (materialise, BC launch, stencil
    and product kernels) against the host path and the numpy restatement."""
    from oracle import dsl
    run([BIN / "composite", "cpu", tmp_path, mode])
    run([BIN / "composite", "gpu", tmp_path, mode])
    for name, want in dsl.composite_forms(*dsl.composite_inputs(), 0.5, 0.25, mode).items():
        _, h = read_field(tmp_path / f"{name}_{mode}_cpu.field")
        _, d = read_field(tmp_path / f"{name}_{mode}_gpu.field")
        scale = max(1.0, np.abs(want).max())
        e_host, e_ref = np.abs(d - h).max(), np.abs(d[0, 0] - want).max()
        print(f"{mode:8s} {name:10s} device vs host {e_host:.2e}, vs numpy {e_ref:.2e}")
        assert e_host <= 1e-12 * scale and e_ref <= 1e-12 * scale, (name, e_host, e_ref)
