"""Learned surrogate for 2D periodic Cahn-Hilliard, judged by the solver.

Host orchestration over the C ABI (include/phix_amd.h); torch provides device
memory and the stream only. Pieces:

  * `generate_pairs` -- training data from the solver itself: an ensemble of
    256^2 runs advanced by the fused FP32 RK4 ensemble kernel, snapshots every
    `every` steps, pairs (c(t), c(t + every*dt)).
  * `Surrogate` -- the 6-layer BF16 MFMA conv net (phx_sur_*), Adam training.
  * `HybridSolver` -- surrogate rollout in blocks of `check_every` surrogate
    steps; after each block the solver judges it: the free energy (FP64
    reduction) must not rise, and one surrogate step is re-run by the solver
    (`every` RK4 steps) whose result must agree within `tol` (relative L2).
    A rejected block is recomputed by the solver -- the fallback.
  * `PlainSolver` -- the same problem on the fused RK4 solver only.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi, init


def _dev_empty(n, dtype, dev):
    import torch
    return torch.empty(n, dtype=dtype, device=dev)


# ---------------------------------------------------------------- solver ----
class PlainSolver:
    """Fused RK4 solver for B periodic n x n fields (f = c^2(1-c)^2, M = 1)."""

    def __init__(self, dev, n=256, kappa=0.5, dt=0.01, dtype=capi.F32):
        import torch
        self.dev, self.n, self.dt, self.dtype = dev, n, dt, dtype
        self.g = capi.make_grid((n, n), bc=capi.PERIODIC)
        self.p = capi.CHParams(1.0, 0.0, 1.0, kappa, 1.0, dt)
        self.tdt = torch.float32 if dtype == capi.F32 else torch.float64
        esz = 4 if dtype == capi.F32 else 8
        self.shape = capi.stored_shape(self.g)
        self.c = torch.zeros(self.shape, dtype=self.tdt, device=dev)
        self.work = torch.empty(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(self.g), dtype) // esz,
                                dtype=self.tdt, device=dev)
        self.red = torch.empty(capi.lib().phx_reduce_workspace_bytes(C.byref(self.g)) // 8, dtype=torch.float64,
                               device=dev)
        self.out = torch.empty(2, dtype=torch.float64, device=dev)

    def phys(self):
        return self.c[1, 1:self.n + 1, 1:self.n + 1]

    def set(self, c2d):
        self.phys().copy_(c2d.to(self.tdt))

    def steps(self, k):
        if k > 0:
            capi.call("phx_ch_rk4_steps", C.byref(self.g), self.dtype, capi.ptr(self.c), capi.ptr(self.work),
                      C.byref(self.p), int(k), capi.stream_handle())

    def mass_energy(self):
        capi.call("phx_mass_energy", C.byref(self.g), self.dtype, capi.ptr(self.c), C.byref(self.p),
                  capi.ptr(self.red), capi.ptr(self.out), capi.stream_handle())
        m, e = self.out.tolist()
        return m, e


# ---------------------------------------------------------------- data ------
def generate_pairs(dev, ntraj, nsnap, seed, n=256, kappa=0.5, dt=0.01, every=100, amp=0.05, skip=0):
    """Returns (inputs, targets): (ntraj*nsnap, n, n) FP32 device tensors of
    consecutive snapshots c(t_k), c(t_k + every*dt), t_k = (skip + k)*every*dt,
    from c0 = 0.5 + U(-amp, amp) (SplitMix64 substream per trajectory)."""
    import torch
    h = C.c_void_p()
    sides = (C.c_int * ntraj)(*([n] * ntraj))
    kap = (C.c_double * ntraj)(*([kappa] * ntraj))
    capi.call("phx_ensemble_create", ntraj, sides, kap, None, C.byref(h))
    try:
        cells = capi.lib().phx_ensemble_cells(h)
        x0 = np.concatenate([init.uniform(init.substream(seed, 100 + r), n * n, 0.5 - amp, 0.5 + amp)
                             for r in range(ntraj)])
        c = torch.from_numpy(x0).float().to(dev)
        work = torch.empty(3 * cells, dtype=torch.float32, device=dev)
        snaps = torch.empty((nsnap + 1, ntraj, n, n), dtype=torch.float32, device=dev)
        sh = capi.stream_handle()
        if skip:
            capi.call("phx_ensemble_graph_steps", h, capi.ptr(c), capi.ptr(work), dt, every, skip, sh)
        snaps[0] = c.view(ntraj, n, n)
        for k in range(nsnap):
            capi.call("phx_ensemble_graph_steps", h, capi.ptr(c), capi.ptr(work), dt, every, 1, sh)
            snaps[k + 1] = c.view(ntraj, n, n)
        inputs = snaps[:-1].reshape(-1, n, n).contiguous()
        targets = snaps[1:].reshape(-1, n, n).contiguous()
        return inputs, targets
    finally:
        capi.lib().phx_ensemble_destroy(h)


# -------------------------------------------------------------- network -----
class Surrogate:
    def __init__(self, dev, batch, n=256, seed=0, params=None):
        import torch
        self.dev, self.B, self.n = dev, batch, n
        self.npar = capi.lib().phx_sur_param_count()
        if params is None:
            params = self.init_params(seed)
        self.p = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(dev)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.grad = torch.zeros_like(self.p)
        self.ws = torch.empty(capi.lib().phx_sur_workspace_bytes(batch, n, n), dtype=torch.uint8, device=dev)
        self.out = torch.empty((batch, n, n), dtype=torch.float32, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
        self.step = 0

    def init_params(self, seed):
        """He-normal conv weights, zero biases, zero last layer (the initial
        network is the identity map c' = c)."""
        rng = np.random.default_rng(seed)
        parts = []
        for i, (co, ci) in enumerate([(32, 1), (32, 32), (32, 32), (32, 32), (32, 32), (1, 32)]):
            std = math.sqrt(2.0 / (ci * 9))
            w = rng.standard_normal(co * ci * 9) * std if i < 5 else np.zeros(co * ci * 9)
            parts += [w, np.zeros(co)]
        return np.concatenate(parts).astype(np.float32)

    def forward(self, c, reuse_weights=False):
        assert c.shape == (self.B, self.n, self.n) and c.dtype.itemsize == 4 and c.is_contiguous()
        capi.call("phx_sur_forward", capi.ptr(self.p), capi.ptr(self.ws), capi.ptr(c), capi.ptr(self.out), self.B,
                  self.n, self.n, 1 if reuse_weights else 0, capi.stream_handle())
        return self.out

    def train_step(self, c, tgt, lr):
        self.forward(c)
        capi.call("phx_sur_backward", capi.ptr(self.p), capi.ptr(self.ws), capi.ptr(c), capi.ptr(self.out),
                  capi.ptr(tgt), capi.ptr(self.grad), capi.ptr(self.loss), self.B, self.n, self.n,
                  capi.stream_handle())
        self.step += 1
        capi.call("phx_sur_adam", capi.ptr(self.p), capi.ptr(self.grad), capi.ptr(self.m), capi.ptr(self.v),
                  self.step, lr, 0.9, 0.999, 1e-8, capi.stream_handle())
        return self.loss

    def params_numpy(self):
        return self.p.detach().cpu().numpy()


def train(dev, inputs, targets, batch=8, epochs=1, lr=1e-3, seed=0, params=None, log=None, max_seconds=None):
    """Adam on shuffled pairs; returns (Surrogate, list of mean losses per epoch)."""
    import torch
    net = Surrogate(dev, batch, n=inputs.shape[-1], seed=seed, params=params)
    npairs = inputs.shape[0]
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(npairs)).to(dev)
        tot, cnt = 0.0, 0
        acc = torch.zeros(1, dtype=torch.float64, device=dev)
        for b0 in range(0, npairs - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            x = inputs.index_select(0, idx).contiguous()
            y = targets.index_select(0, idx).contiguous()
            lr_now = lr * 0.5 * (1 + math.cos(math.pi * (ep * npairs + b0) / (epochs * npairs)))
            acc += net.train_step(x, y, lr_now)
            cnt += 1
            if max_seconds and time.time() - t0 > max_seconds:
                break
        tot = acc.item() / (cnt * batch * inputs.shape[-1] ** 2)
        hist.append(tot)
        if log:
            log(f"epoch {ep}: mse {tot:.3e} ({time.time() - t0:.1f} s)")
        if max_seconds and time.time() - t0 > max_seconds:
            break
    return net, hist


# --------------------------------------------------------------- hybrid -----
@dataclass
class HybridReport:
    seconds: float
    blocks: int
    rejected: int
    surrogate_steps: int
    solver_steps: int
    log: list = field(default_factory=list)


class HybridSolver:
    """Surrogate rollout judged by the solver, with solver fallback."""

    def __init__(self, dev, params, n=256, kappa=0.5, dt=0.01, every=100, check_every=10, tol=2e-3, etol=1e-6):
        self.net = Surrogate(dev, 1, n=n, params=params)
        self.solver = PlainSolver(dev, n=n, kappa=kappa, dt=dt)
        self.every, self.check_every, self.tol, self.etol = every, check_every, tol, etol
        self.n = n
        self.packed = False

    def run(self, c0, nsteps):
        """Advances c0 (n x n FP32 device tensor) by nsteps solver steps (a
        multiple of `every`); returns (final field, HybridReport)."""
        import torch
        assert nsteps % self.every == 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cur = c0.reshape(1, self.n, self.n).contiguous().clone()
        remaining = nsteps // self.every
        blocks = rejected = ssteps = vsteps = 0
        S = self.solver
        S.set(cur[0])
        _, e_prev = S.mass_energy()
        log = []
        while remaining > 0:
            k = min(self.check_every, remaining)
            start = cur.clone()
            prev = cur
            for _ in range(k):
                prev = cur
                cur = self.net.forward(cur, reuse_weights=self.packed).clone()
                self.packed = True
            ssteps += k
            # judge: one surrogate step against `every` solver steps from its input
            S.set(prev[0])
            S.steps(self.every)
            vsteps += self.every
            ref = S.phys()
            rel = ((cur[0] - ref).norm() / ref.norm()).item()
            S.set(cur[0])
            _, e_now = S.mass_energy()
            ok = rel <= self.tol and e_now <= e_prev * (1 + self.etol)
            if not ok:
                rejected += 1
                S.set(start[0])
                S.steps(k * self.every)
                vsteps += k * self.every
                cur = S.phys().float().reshape(1, self.n, self.n).contiguous().clone()
                _, e_now = S.mass_energy()
            log.append((blocks, rel, e_now, ok))
            e_prev = e_now
            blocks += 1
            remaining -= k
        torch.cuda.synchronize()
        return cur[0], HybridReport(time.perf_counter() - t0, blocks, rejected, ssteps, vsteps, log)


# ------------------------------------------------------------------ CLI -----
DATA = __import__("pathlib").Path(__file__).resolve().parent / "data"
WEIGHTS = DATA / "surrogate_ch256.npz"


def fp64_reference(dev, c0_np, nsteps, n=256, kappa=0.5, dt=0.01):
    """FP64 fused solver from c0 (numpy n x n) for nsteps; returns a numpy field."""
    import torch
    S = PlainSolver(dev, n=n, kappa=kappa, dt=dt, dtype=capi.F64)
    S.set(torch.from_numpy(c0_np).to(dev))
    S.steps(nsteps)
    torch.cuda.synchronize()
    return S.phys().cpu().numpy()


def main(argv=None):
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser(description="train or evaluate the CH surrogate")
    ap.add_argument("mode", choices=["train", "eval"])
    ap.add_argument("--ntraj", type=int, default=64)
    ap.add_argument("--nsnap", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seconds", type=float, default=600)
    ap.add_argument("--tend-steps", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--tol", type=float, default=5e-3)
    ap.add_argument("--check-every", type=int, default=10)
    a = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    if a.mode == "train":
        t0 = time.time()
        xs, ys = generate_pairs(dev, a.ntraj, a.nsnap, seed=a.seed)
        torch.cuda.synchronize()
        print(f"data: {xs.shape[0]} pairs in {time.time() - t0:.1f} s", flush=True)
        net, hist = train(dev, xs, ys, batch=a.batch, epochs=a.epochs, lr=a.lr, log=lambda s: print(s, flush=True),
                          max_seconds=a.seconds)
        DATA.mkdir(exist_ok=True)
        np.savez(a.out, params=net.params_numpy(), mse=np.array(hist), ntraj=a.ntraj, nsnap=a.nsnap,
                 seed=a.seed)
        print(json.dumps({"saved": a.out, "mse": hist}), flush=True)
    else:
        p = np.load(a.out)["params"]
        c0 = init.uniform(init.substream(1234, 7), 256 * 256, 0.45, 0.55).reshape(256, 256)
        ref = fp64_reference(dev, c0, a.tend_steps)
        hyb = HybridSolver(dev, p, tol=a.tol, check_every=a.check_every)
        hyb.run(torch.from_numpy(c0).float().to(dev), 10 * hyb.every)   # warm-up (code paths, weight packing)
        out, rep = hyb.run(torch.from_numpy(c0).float().to(dev), a.tend_steps)
        for b in rep.log:
            print(f"block {b[0]}: one-step residual {b[1]:.3e} energy {b[2]:.6e} accepted {b[3]}", flush=True)
        err = float(np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref))
        plain = PlainSolver(dev)
        plain.set(torch.from_numpy(c0).float().to(dev))
        plain.steps(100)
        plain.set(torch.from_numpy(c0).float().to(dev))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain.steps(a.tend_steps)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        perr = float(np.linalg.norm(plain.phys().cpu().numpy() - ref) / np.linalg.norm(ref))
        print(json.dumps({"t_end_steps": a.tend_steps, "tol": a.tol, "hybrid_s": rep.seconds, "hybrid_rel_l2": err,
                          "rejected": rep.rejected, "blocks": rep.blocks, "surrogate_steps": rep.surrogate_steps,
                          "solver_steps": rep.solver_steps, "plain_s": tp, "plain_rel_l2": perr,
                          "speedup": tp / rep.seconds}), flush=True)


if __name__ == "__main__":
    main()
