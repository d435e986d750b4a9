"""Batched uneven-size ensemble (config 4 family): every run of the packed
FP32 ensemble against its own FP64 oracle run; hipGraph replay == eager
launches bit for bit; tiling choices (autotuner candidates) do not change
results; per-run FP64 mass is conserved."""
import ctypes as C

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, stream

pytestmark = pytest.mark.gpu


class Ens:
    def __init__(self, sides, kappa):
        self.h = C.c_void_p()
        n = len(sides)
        capi.call("phx_ensemble_create", n, (C.c_int * n)(*sides), (C.c_double * n)(*kappa), None, C.byref(self.h))
        self.cells = capi.lib().phx_ensemble_cells(self.h)
        self.off = [capi.lib().phx_ensemble_offset(self.h, r) for r in range(n)]

    def close(self):
        capi.lib().phx_ensemble_destroy(self.h)


def make_state(sides, cbar, seed):
    parts = []
    for r, n in enumerate(sides):
        parts.append(init.uniform(init.substream(seed, 10 + r), n * n, cbar[r] - 0.01, cbar[r] + 0.01))
    return np.concatenate(parts)


def test_ensemble_matches_oracle_per_run(dev):
    import torch
    sides = [64, 40, 96, 33, 128, 2 * 32 + 64]
    kappa = [0.25, 0.5, 0.75, 1.0, 0.6, 0.3]
    cbar = [0.3, 0.4, 0.5, 0.6, 0.7, 0.45]
    E = Ens(sides, kappa)
    x0 = make_state(sides, cbar, 4)
    c = torch.from_numpy(x0).float().to(dev)
    work = torch.empty(3 * E.cells, dtype=torch.float32, device=dev)
    steps = 100
    capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(c), capi.ptr(work), 0.01, steps, stream())
    got = c.cpu().double().numpy()
    worst = 0.0
    for r, n in enumerate(sides):
        g = capi.make_grid((n, n), bc=capi.PERIODIC)
        phys = x0[E.off[r]:E.off[r] + n * n].astype(np.float32).astype(np.float64).reshape(1, n, n)
        a = init.to_stored(phys, g)
        O.ch_rk4_steps(g, ch_params(kappa=kappa[r], dt=0.01), a, steps)
        err = np.abs(got[E.off[r]:E.off[r] + n * n].reshape(1, n, n) - capi.physical(a, g)).max()
        worst = max(worst, err)
    print(f"ensemble fp32 vs fp64 oracle after {steps} steps: worst max-abs {worst:.3e}")
    assert worst <= 2e-6
    E.close()


def test_graph_replay_equals_eager_and_tiling_invariance(dev):
    import torch
    sides, kappa = [64, 96, 160, 32], [0.3, 0.6, 0.9, 0.5]
    E = Ens(sides, kappa)
    x0 = torch.from_numpy(make_state(sides, [0.4, 0.5, 0.6, 0.5], 5)).float().to(dev)
    work = torch.empty(3 * E.cells, dtype=torch.float32, device=dev)
    a = x0.clone()
    capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(a), capi.ptr(work), 0.01, 30, stream())
    b = x0.clone()
    for _ in range(3):
        capi.call("phx_ensemble_graph_steps", E.h, capi.ptr(b), capi.ptr(work), 0.01, 5, 2, stream())
        torch.cuda.synchronize()
    assert torch.equal(a, b)
    for shape in range(5):
        capi.call("phx_ensemble_set_tiling", E.h, shape, 37)
        t = x0.clone()
        capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(t), capi.ptr(work), 0.01, 30, stream())
        assert torch.equal(a, t), shape
    scratch = x0.clone()
    ms = C.c_double()
    capi.call("phx_ensemble_autotune", E.h, capi.ptr(scratch), capi.ptr(work), 0.01, 3, C.byref(ms), stream())
    sh, gr, tx, ty = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    capi.call("phx_ensemble_get_tiling", E.h, C.byref(sh), C.byref(gr), C.byref(tx), C.byref(ty))
    assert 0 <= sh.value < 5 and gr.value > 0 and ms.value > 0
    E.close()


def test_config4_full_ensemble_mass(dev):
    """Config 4 at its full size: 512 runs, N = 64 + 32*floor(15u), kappa ~ U(0.25,1),
    cbar ~ U(0.3,0.7) +- 0.01, RK4 dt=0.01, 1000 steps through the captured graph;
    per-run FP64 mass is conserved to FP32 round-off."""
    import torch
    N, kappa, cbar = init.ensemble_params(512, 4)
    E = Ens([int(v) for v in N], [float(v) for v in kappa])
    x0 = make_state([int(v) for v in N], cbar, 4)
    c = torch.from_numpy(x0).float().to(dev)
    work = torch.empty(3 * E.cells, dtype=torch.float32, device=dev)
    m0 = torch.empty(512, dtype=torch.float64, device=dev)
    m1 = torch.empty(512, dtype=torch.float64, device=dev)
    capi.call("phx_ensemble_mass", E.h, capi.ptr(c), capi.ptr(m0), stream())
    capi.call("phx_ensemble_graph_steps", E.h, capi.ptr(c), capi.ptr(work), 0.01, 100, 10, stream())
    capi.call("phx_ensemble_mass", E.h, capi.ptr(c), capi.ptr(m1), stream())
    torch.cuda.synchronize()
    rel = ((m1 - m0).abs() / m0.abs()).max().item()
    print(f"config4: {E.cells} cells, max per-run relative mass drift after 1000 steps {rel:.3e}")
    assert torch.isfinite(c).all()
    assert rel < 1e-5
    # mean compositions stay at the drawn cbar
    means = (m1.cpu().numpy() / (N.astype(np.float64) ** 2))
    assert np.abs(means - cbar).max() < 0.02
    E.close()


@pytest.mark.parametrize("steps", [1, 2, 7])
def test_two_stage_equals_one_stage(dev, steps):
    """The two-stage kernel (stages 1+2 and 3+4 per launch, one workgroup per
    tile) returns the persistent one-stage kernel's bits for every tile shape,
    odd and even step counts (the state alternates between c and the
    workspace), tiny runs (sides 2, 3, 5: halos wrap more than once) and the
    captured graph."""
    import torch
    sides, kappa = [2, 3, 5, 64, 96, 33, 130], [0.3, 0.9, 0.5, 0.6, 0.25, 1.0, 0.7]
    E = Ens(sides, kappa)
    x0 = torch.from_numpy(make_state(sides, [0.5, 0.4, 0.6, 0.5, 0.45, 0.55, 0.35], 6)).float().to(dev)
    work = torch.empty(3 * E.cells, dtype=torch.float32, device=dev)
    capi.call("phx_ensemble_set_two_stage", E.h, 0)
    capi.call("phx_ensemble_set_tiling", E.h, 0, 0)
    ref = x0.clone()
    capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(ref), capi.ptr(work), 0.01, steps, stream())
    capi.call("phx_ensemble_set_two_stage", E.h, 1)
    assert capi.lib().phx_ensemble_get_two_stage(E.h) == 1
    for shape in range(5):
        capi.call("phx_ensemble_set_tiling", E.h, shape, 0)
        t = x0.clone()
        capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(t), capi.ptr(work), 0.01, steps, stream())
        torch.cuda.synchronize()
        print(f"steps {steps} shape {shape}: max |two - one| {(t - ref).abs().max().item():.3e}", flush=True)
        assert torch.equal(t, ref), shape
    g = x0.clone()
    capi.call("phx_ensemble_graph_steps", E.h, capi.ptr(g), capi.ptr(work), 0.01, steps, 2, stream())
    r2 = ref.clone()
    capi.call("phx_ensemble_set_two_stage", E.h, 0)
    capi.call("phx_ensemble_rk4_steps", E.h, capi.ptr(r2), capi.ptr(work), 0.01, steps, stream())
    torch.cuda.synchronize()
    assert torch.equal(g, r2)
    assert not torch.equal(ref, x0)
    E.close()
