"""GPU parity of the fused Cahn-Hilliard kernels against the FP64 oracle.

Contract: the FP64 GPU path agrees with the FP64 oracle within 1e-12 max-abs
at every step (it is designed to agree bit for bit; the max error is printed).
FP32 bulk: its error against the FP64 path is stated per configuration."""
import ctypes as C
import math

import numpy as np
import pytest

from phix_amd import capi, init
from oracle import oracle as O

from helpers import ch_params, to_dev, to_np, stream

pytestmark = pytest.mark.gpu
P, N, F = capi.PERIODIC, capi.NOFLUX, capi.FIXED
TOL64 = 1e-12


def test_config1_euler_every_step(dev):
    """Config 1: 128^2, dx=1, periodic, c0 = 0.5+U(-0.01,0.01) seed 1, kappa=0.5,
    M=1, dt=0.01, multi-step Euler (mu STEADY, c TRANSIENT), 1000 steps, FP64.
    Every stored cell of the active region (physical + ghost ring) of c and mu
    is compared after every step."""
    n = (128, 128)
    g = capi.make_grid(n, bc=P)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=0.5, M=1.0, dt=0.01)
    c = init.to_stored(init.ch_field(n, 1, 0.5, 0.01), g)
    mu = np.zeros_like(c)
    tc, tm = to_dev(c, dev), to_dev(mu, dev)
    worst = 0.0
    for s in range(1000):
        capi.call("phx_ch_euler_steps", C.byref(g), capi.F64, capi.ptr(tc), capi.ptr(tm), C.byref(p), 1, stream())
        O.ch_euler_steps(g, p, c, mu, 1)
        dc = np.abs(capi.active(to_np(tc), g) - capi.active(c, g)).max()
        dm = np.abs(capi.active(to_np(tm), g) - capi.active(mu, g)).max()
        worst = max(worst, dc, dm)
        assert dc <= TOL64 and dm <= TOL64, (s, dc, dm)
    print(f"config1 euler: worst per-step max-abs over 1000 steps = {worst:.3e}")
    # outputs: mean and total energy
    m, e = O.mass_energy(g, p, c)
    assert abs(m / (128 * 128) - 0.5) < 1e-3


def test_config1_euler_one_call_against_oracle(dev):
    """Config 1 as the CH app advances it: all 1000 steps in one
    phx_ch_euler_steps call (249 launches of four steps, three of one, the
    last step with the two kernels), every stored cell of the active region
    of c and mu against the FP64 oracle's 1000 steps: equal bit for bit."""
    n = (128, 128)
    g = capi.make_grid(n, bc=P)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=0.5, M=1.0, dt=0.01)
    c = init.to_stored(init.ch_field(n, 1, 0.5, 0.01), g)
    mu = np.zeros_like(c)
    tc, tm = to_dev(c, dev), to_dev(mu, dev)
    capi.call("phx_ch_euler_steps", C.byref(g), capi.F64, capi.ptr(tc), capi.ptr(tm), C.byref(p), 1000, stream())
    O.ch_euler_steps(g, p, c, mu, 1000)
    assert np.array_equal(capi.active(to_np(tc), g), capi.active(c, g))
    assert np.array_equal(capi.active(to_np(tm), g), capi.active(mu, g))


def test_euler_multistep_batched_matches_single(dev):
    g = capi.make_grid((33, 17, 9), bc=[[N, N], [F, N], [P, P]], bc_value=[[0, 0], [0.4, 0], [0, 0]])
    p = ch_params(rho=1.0, kappa=0.5, dt=0.005)
    c = init.to_stored(init.ch_field((33, 17, 9), 5, 0.45, 0.05), g)
    mu = np.zeros_like(c)
    tc, tm = to_dev(c, dev), to_dev(mu, dev)
    capi.call("phx_ch_euler_steps", C.byref(g), capi.F64, capi.ptr(tc), capi.ptr(tm), C.byref(p), 37, stream())
    O.ch_euler_steps(g, p, c, mu, 37)
    assert np.abs(capi.active(to_np(tc), g) - capi.active(c, g)).max() <= TOL64
    assert np.abs(capi.active(to_np(tm), g) - capi.active(mu, g)).max() <= TOL64


RK4_CASES = [
    # n, d, ghost, bc, values
    ((64, 48), (1.0, 1.0), 1, P, 0.0),
    ((37, 29), (0.9, 1.2), 2, [[N, N], [N, N]], 0.0),
    ((40, 33), (1.0, 1.0), 1, [[F, N], [P, P]], [[0.45, 0], [0, 0]]),
    ((5,) * 0 + (70,), (1.0,), 1, N, 0.0),
    ((40, 36, 34), (1.0, 1.0, 1.0), 1, N, 0.0),
    ((67, 20, 35), (1.0, 0.8, 1.1), 1, [[P, P], [N, F], [F, N]], [[0, 0], [0, 0.42], [0.38, 0]]),
    ((32, 24, 70), (1.0, 1.0, 1.0), 2, P, 0.0),
]


@pytest.mark.parametrize("case", range(len(RK4_CASES)))
def test_rk4_fp64_every_step(dev, case):
    import torch
    n, d, ghost, bc, vals = RK4_CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    tc = to_dev(c, dev)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F64) // 8, dtype=torch.float64,
                       device=dev)
    ow = np.zeros((6,) + c.shape)
    steps = 60 if len(n) == 3 else 200
    worst = 0.0
    for s in range(steps):
        capi.call("phx_ch_rk4_steps", C.byref(g), capi.F64, capi.ptr(tc), capi.ptr(work), C.byref(p), 1, stream())
        O.ch_rk4_steps(g, p, c, 1, ow)
        err = np.abs(capi.active(to_np(tc), g) - capi.active(c, g)).max()
        worst = max(worst, err)
        assert err <= TOL64, (s, err)
    print(f"rk4 fp64 case {case}: worst per-step max-abs = {worst:.3e}")


@pytest.mark.parametrize("case", [0, 2, 4, 5])
def test_rk4_fp32_error_bound(dev, case):
    """FP32 bulk against the FP64 oracle: max-abs after 100 steps <= 2e-6."""
    import torch
    n, d, ghost, bc, vals = RK4_CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, kappa=1.0, dt=0.005)
    c = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    tc = to_dev(c, dev, torch.float32)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F32) // 4, dtype=torch.float32,
                       device=dev)
    capi.call("phx_ch_rk4_steps", C.byref(g), capi.F32, capi.ptr(tc), capi.ptr(work), C.byref(p), 100, stream())
    O.ch_rk4_steps(g, p, c, 100)
    err = np.abs(capi.physical(to_np(tc), g) - capi.physical(c, g)).max()
    print(f"rk4 fp32 case {case}: max-abs after 100 steps = {err:.3e}")
    assert err <= 2e-6


def test_rk4_multistep_call_equals_single_steps(dev):
    import torch
    g = capi.make_grid((48, 40, 24), bc=N)
    p = ch_params(kappa=1.0, dt=0.005)
    c = init.to_stored(init.ch_field((48, 40, 24), 7, 0.4, 0.02), g)
    ta, tb = to_dev(c, dev), to_dev(c, dev)
    work = torch.zeros(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F64) // 8, dtype=torch.float64,
                       device=dev)
    capi.call("phx_ch_rk4_steps", C.byref(g), capi.F64, capi.ptr(ta), capi.ptr(work), C.byref(p), 25, stream())
    for _ in range(25):
        capi.call("phx_ch_rk4_steps", C.byref(g), capi.F64, capi.ptr(tb), capi.ptr(work), C.byref(p), 1, stream())
    assert torch.equal(ta, tb)


def test_config2_full_size_properties(dev):
    """Config 2 at its full size: 256^3, NoFlux, c0 = 0.4+U(-0.02,0.02) seed 2,
    kappa=1, M=1, dt=0.005, RK4, 500 steps, bulk FP64 and FP32 run in lockstep.
    The FP64 path (bit-exact with the oracle at the sizes above) stands in as the
    full-size FP64 reference; per-step FP32 max-abs and the final relative L2 are
    reported and bounded. Mass (FP64 compensated reduction) is conserved."""
    import torch
    n = (256, 256, 256)
    g = capi.make_grid(n, bc=N)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=1.0, M=1.0, dt=0.005)
    c0 = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    t64 = to_dev(c0, dev)
    t32 = to_dev(c0, dev, torch.float32)
    del c0
    w64 = torch.empty(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F64) // 8, dtype=torch.float64, device=dev)
    w32 = torch.empty(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), capi.F32) // 4, dtype=torch.float32, device=dev)
    red = torch.empty(capi.lib().phx_reduce_workspace_bytes(C.byref(g)) // 8, dtype=torch.float64, device=dev)
    out = torch.empty(2, dtype=torch.float64, device=dev)
    capi.call("phx_mass_energy", C.byref(g), capi.F64, capi.ptr(t64), C.byref(p), capi.ptr(red), capi.ptr(out), stream())
    m0, e0 = out.tolist()
    gh = 1
    sl = (slice(gh, gh + 256),) * 3
    worst = 0.0
    for s in range(500):
        capi.call("phx_ch_rk4_steps", C.byref(g), capi.F64, capi.ptr(t64), capi.ptr(w64), C.byref(p), 1, stream())
        capi.call("phx_ch_rk4_steps", C.byref(g), capi.F32, capi.ptr(t32), capi.ptr(w32), C.byref(p), 1, stream())
        if s % 10 == 9:
            worst = max(worst, (t64.view(258, 258, 258)[sl] - t32.view(258, 258, 258)[sl].double()).abs().max().item())
    a = t64.view(258, 258, 258)[sl]
    b = t32.view(258, 258, 258)[sl].double()
    rel = ((a - b).norm() / a.norm()).item()
    capi.call("phx_mass_energy", C.byref(g), capi.F64, capi.ptr(t64), C.byref(p), capi.ptr(red), capi.ptr(out), stream())
    m1, e1 = out.tolist()
    capi.call("phx_mass_energy", C.byref(g), capi.F32, capi.ptr(t32), C.byref(p), capi.ptr(red), capi.ptr(out), stream())
    m32, e32 = out.tolist()
    print(f"config2 256^3 RK4 500 steps: fp32 vs fp64 max-abs (every 10th step) {worst:.3e}, final rel-L2 {rel:.3e}; "
          f"mass fp64 drift {abs(m1 - m0) / m0:.3e}, fp32 drift {abs(m32 - m0) / m0:.3e}; energy {e0:.6e} -> {e1:.6e}")
    assert abs(m1 - m0) <= 1e-12 * abs(m0)
    assert abs(m32 - m0) <= 1e-5 * abs(m0)
    assert worst <= 1e-5 and rel <= 1e-5
    assert e1 < e0


@pytest.mark.parametrize("dtype", [capi.F64, capi.F32])
@pytest.mark.parametrize("case", [0, 1, 2, 4, 5])
def test_mass_energy_against_oracle(dev, dtype, case):
    import torch
    n, d, ghost, bc, vals = RK4_CASES[case]
    g = capi.make_grid(n, d=d, ghost=ghost, bc=bc, bc_value=vals)
    p = ch_params(rho=2.0, ca=0.2, cb=0.8, kappa=0.7)
    c = init.to_stored(init.ch_field(n, 11, 0.45, 0.3), g)
    if dtype == capi.F32:
        c = c.astype(np.float32).astype(np.float64)
    tc = to_dev(c, dev, torch.float64 if dtype == capi.F64 else torch.float32)
    red = torch.empty(capi.lib().phx_reduce_workspace_bytes(C.byref(g)) // 8, dtype=torch.float64, device=dev)
    out = torch.empty(2, dtype=torch.float64, device=dev)
    capi.call("phx_mass_energy", C.byref(g), dtype, capi.ptr(tc), C.byref(p), capi.ptr(red), capi.ptr(out), stream())
    m, e = out.tolist()
    mo, eo = O.mass_energy(g, p, c)
    assert abs(m - mo) <= 1e-14 * abs(mo)
    assert abs(e - eo) <= 1e-13 * abs(eo)
    # determinism
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    capi.call("phx_mass_energy", C.byref(g), dtype, capi.ptr(tc), C.byref(p), capi.ptr(red), capi.ptr(out2), stream())
    assert torch.equal(out, out2)


@pytest.mark.parametrize("dtype", [capi.F64, capi.F32])
@pytest.mark.parametrize("n,bc,vals,nsteps", [
    ((128, 128), P, 0.0, 2),
    ((128, 128), P, 0.0, 37),
    ((70, 33), N, 0.0, 10),
    ((41, 30), [[F, N], [P, P]], [[0.3, 0], [0, 0]], 5),
    ((30, 41), [[P, P], [N, F]], [[0, 0], [0, 0.6]], 4),
    ((3, 2), N, 0.0, 3),
    ((50,), P, 0.0, 9),
    ((32, 32, 32), P, 0.0, 6),
    ((70, 20, 37), N, 0.0, 5),
    ((40, 9, 33), [[F, N], [P, P], [N, F]], [[0.3, 0], [0, 0], [0, 0.6]], 4),
    ((9, 17, 2), [[P, P], [F, F], [P, P]], [[0, 0], [0.45, 0.55], [0, 0]], 3),
    ((2, 3, 18), [[N, N], [N, N], [F, N]], [[0, 0], [0, 0], [0.4, 0]], 7),
])
def test_euler_one_launch_per_step_equals_two(dev, dtype, n, bc, vals, nsteps):
    """Euler in 1D, 2D and 3D: the one-launch-per-step path returns the
    two-kernel path's bits for c and mu (ghost layers included), odd and even
    step counts, partial tiles and several z-chunks; FP64 also against the
    oracle."""
    g = capi.make_grid(n, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=0.5, M=1.0, dt=0.01)
    c0 = init.to_stored(init.ch_field(n, 9, 0.5, 0.05), g)
    out = {}
    for fused in (0, 1):
        capi.call("phx_ch_euler_set_fused", fused)
        try:
            tc = to_dev(c0, dev) if dtype == capi.F64 else to_dev(c0.astype(np.float32), dev)
            tm = torch_zeros_like(tc)
            capi.call("phx_ch_euler_steps", C.byref(g), dtype, capi.ptr(tc), capi.ptr(tm), C.byref(p), nsteps,
                      stream())
            out[fused] = (to_np(tc), to_np(tm))
        finally:
            capi.call("phx_ch_euler_set_fused", 1)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    if dtype == capi.F64:
        c, mu = c0.copy(), np.zeros_like(c0)
        O.ch_euler_steps(g, p, c, mu, nsteps)
        assert np.abs(capi.active(out[1][0], g) - capi.active(c, g)).max() <= TOL64
        assert np.abs(capi.active(out[1][1], g) - capi.active(mu, g)).max() <= TOL64


def torch_zeros_like(t):
    import torch
    return torch.zeros_like(t)


@pytest.mark.parametrize("dtype", [capi.F64, capi.F32])
@pytest.mark.parametrize("n,bc,vals", [
    ((128, 128), P, 0.0),
    ((72, 40), P, 0.0),
    ((8, 8), P, 0.0),
    ((9, 21), P, 0.0),
    ((70, 33), N, 0.0),
    ((41, 30), [[F, N], [P, P]], [[0.3, 0], [0, 0]]),
    ((30, 41), [[P, P], [N, F]], [[0, 0], [0, 0.6]]),
    ((8, 12), [[F, F], [N, N]], [[0.45, 0.55], [0, 0]]),
])
def test_euler_four_steps_per_launch_equals_two_kernels(dev, dtype, n, bc, vals):
    """2D Euler in blocks of four steps per launch (every tile of the
    whole-step kernel's Euler form, every BC kind, step counts that leave 0-3
    single steps and odd or even launch counts) returns the two-kernel
    path's bits for c and mu, ghost layers included."""
    g = capi.make_grid(n, bc=bc, bc_value=vals)
    p = ch_params(rho=1.0, ca=0.0, cb=1.0, kappa=0.5, M=1.0, dt=0.01)
    c0 = init.to_stored(init.ch_field(n, 11, 0.5, 0.05), g)
    if dtype == capi.F32:
        c0 = c0.astype(np.float32)

    def run(nsteps):
        tc = to_dev(c0, dev)
        tm = torch_zeros_like(tc)
        capi.call("phx_ch_euler_steps", C.byref(g), dtype, capi.ptr(tc), capi.ptr(tm), C.byref(p), nsteps, stream())
        return to_np(tc), to_np(tm)

    try:
        for nsteps in (5, 6, 8, 9, 14, 21):
            capi.call("phx_ch_euler_set_fused", 0)
            ref = run(nsteps)
            capi.call("phx_ch_euler_set_fused", 1)
            for tile in (-1, 0, 1, 2, 3):
                capi.call("phx_ch_euler_set_quad", tile)
                got = run(nsteps)
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (nsteps, tile)
    finally:
        capi.call("phx_ch_euler_set_fused", 1)
        capi.call("phx_ch_euler_set_quad", 0)
    assert capi.lib().phx_ch_euler_set_quad(4) == capi.PHX_ERR_INVALID


@pytest.mark.parametrize("n,bc,dtype", [((256, 256), capi.PERIODIC, capi.F32), ((37, 50), capi.NOFLUX, capi.F64),
                                        ((20, 12, 9), capi.PERIODIC, capi.F64)])
def test_mass_energy_batch_equals_single(dev, n, bc, dtype):
    """phx_mass_energy_batch over fields stored back to back returns each
    field's phx_mass_energy bits."""
    import torch
    nb = 5
    g = capi.make_grid(n, bc=bc)
    p = ch_params()
    tdt = torch.float64 if dtype == capi.F64 else torch.float32
    shape = capi.stored_shape(g)
    rng = np.random.default_rng(len(n) * 100 + n[0])
    c = torch.from_numpy(rng.uniform(0.2, 0.8, (nb,) + tuple(shape))).to(tdt).to(dev)
    wb = capi.lib().phx_reduce_workspace_bytes(C.byref(g))
    red = torch.empty(nb * wb // 8, dtype=torch.float64, device=dev)
    outb = torch.empty(2 * nb, dtype=torch.float64, device=dev)
    capi.call("phx_mass_energy_batch", C.byref(g), dtype, nb, capi.ptr(c), C.byref(p), capi.ptr(red),
              capi.ptr(outb), stream())
    outs = torch.empty(2 * nb, dtype=torch.float64, device=dev)
    for b in range(nb):
        capi.call("phx_mass_energy", C.byref(g), dtype, capi.ptr(c[b]), C.byref(p), capi.ptr(red),
                  capi.ptr(outs[2 * b:]), stream())
    torch.cuda.synchronize()
    assert torch.equal(outb, outs)
