"""Benchmark of the fused phase-field hot path on one MI355X.

Default workload (the headline configuration that fits one GPU): 3D
Cahn-Hilliard, 256^3 cells, dx=1, zero-flux BCs on all faces,
f = c^2(1-c)^2, kappa=1, M=1, dt=0.005, classic RK4, FP64 (the reference's
precision); c0 = 0.4 + U(-0.02, 0.02) from SplitMix64 seed 2. One step = one
RK4 time step over the whole grid with the library's default launch for this
shape (--autotune runs the launch autotuner first; --tile/--zc fix a launch).
Inputs are resident in HBM before the timed region. The same workload in FP32
bulk is reported beside it as `companion` (--no-companion skips it).

Timing: W warm-up steps, then --reps windows (default 5) of exactly K steps,
each bracketed by device synchronisation; `value` and `ms_per_step` come from
the median window, every window is listed under `timing`.

Other workloads (--workload), each a configuration of the same path:
  euler128  128^2 periodic CH, multi-step Euler, FP64, kappa=0.5, dt=0.01
  ch2d1024  1024^2 periodic CH RK4, FP64 line + FP32 companion (same physics)
  ac1024    1024^2 periodic Allen-Cahn, 8 order parameters, RK4, FP32
  ens512    512 periodic CH runs, sides 64 + 32*floor(15u), RK4, FP32,
            one persistent launch per stage, hipGraph-captured step loop
  gfa600    the Fe-B glass-formation program's 600^2 case, FP64
  surrogate 256^2 CH to t_end = 200: surrogate-plus-solver (lead-phase parareal)
            vs the plain GPU solver, speedup and relative L2 per initial condition
  copycal   plain 2 GiB device copy (calibrates the PMC byte counters)

Prints ONE JSON line. `roofline` is computed for the dominant kernel from its
algorithmic bytes per launch divided by its average launch duration inside
the median timed window: the same steps run again with HIP events bracketing
every launch on the launch stream give each launch's share, scaled so that
launches x average never exceed the un-instrumented window (the events
themselves add ~9%). `cpu_baseline` times the FP64 CPU oracle (a restatement
of the reference's per-term CPU path) on one thread on a bounded sample of the
same workload, and the same passes on all host cores (`all_cores`).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# launch tile indices of the 3D CH kernels: one-stage 0-4, two-stage 5-8,
# one-stage with z-neighbours in registers 9-17 (columns, workgroup threads)
TILE_NAMES = ["64x16", "64x8", "128x8", "32x16", "64x4", "two-stage 64x8", "two-stage 32x16", "two-stage 32x8",
              "two-stage 64x4", "zreg 64x16/256", "zreg 128x8/256", "zreg 128x16/512", "zreg 256x8/512",
              "zreg 256x16/1024", "zreg 256x4/256", "zreg 256x8/1024", "zreg 128x16/1024", "zreg 128x8/512"]
# 2D launch tile indices: the one-stage kernel, then the whole-step kernel tiles
TILE2_NAMES = ["64x16 one-stage", "whole-step 64x16", "whole-step 64x32", "whole-step 128x16", "whole-step 32x32",
               "whole-step 32x32/512", "whole-step 64x32/512", "whole-step 128x32/1024 (FP64: 64x32/512)",
               "whole-step 32x64/512"]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def traffic_from_profiles(name: str, launch: str | None = None):
    """HBM bytes per launch of the dominant kernel measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if committed for this
    launch (files with a "per_launch" table hold one entry per launch shape;
    a launch that was not profiled reads None)."""
    f = ROOT / "profiles" / f"{name}_traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
    except (OSError, ValueError):
        return None
    if "per_launch" in d:
        e = d["per_launch"].get(launch) if launch is not None else None
        return e.get("bytes_per_launch") if e else None
    return d.get("bytes_per_launch")


def timed(fn, stream):
    import torch
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1)


def timed_reps(run, steps, reps, stream):
    """`reps` timed windows of exactly `steps` steps each (synchronised on both
    sides); returns (median window ms, all window ms)."""
    import statistics
    ms = [timed(lambda: run(steps), stream) for _ in range(reps)]
    return statistics.median(ms), [round(v, 4) for v in ms]


def kernel_times(run, steps):
    """Per-kind launch durations of the CH RK4 kernels: the same `steps` steps
    as the timed region, run again with every launch bracketed by HIP events on
    its stream (kept out of the timed region: the events add ~9% there).
    Returns {kind: (launches, total_ms)}."""
    from phix_amd import capi
    capi.call("phx_ch_timing", 1)
    run(steps)
    kinds = {}
    for k in range(11):
        n, t = C.c_int(), C.c_double()
        capi.call("phx_ch_timing_read", k, C.byref(n), C.byref(t))
        if n.value:
            kinds[k] = (n.value, t.value)
    capi.call("phx_ch_timing", 0)
    return kinds


def launch_shares(kinds, window_ms):
    """Per-kind average launch duration inside the un-instrumented timed
    window: the event-bracketed durations scaled by window / their sum (never
    up), so that launches x average <= the window. Returns {kind: (count,
    avg_ms, event_avg_ms)}."""
    tot = sum(t for _, t in kinds.values())
    f = min(1.0, window_ms / tot) if tot > 0 else 1.0
    return {k: (n, t * f / n, t / n) for k, (n, t) in kinds.items()}


# Algorithmic bytes follow one fixed accounting per workload: the bytes the
# classic stage-fused RK4 moves per unit of work with mu kept on chip, 16
# field passes per step (read c_n; S1 writes u2 and acc; S2, S3 read u, c_n,
# acc and write u', acc; S4 reads u4, acc and writes c_{n+1}): 64 B per cell-step
# in FP32, 128 B in FP64, x fields; Euler: read c, write c. A launch is
# credited with the units of work it completes (a one-stage launch a quarter
# step, a two-stage launch half a step, a whole-step launch one step), so
# kernels that move fewer bytes for the same work (temporal blocking, the 3D
# stage-state form) can exceed the one-stage bound. The kernel's own
# bytes and the HBM traffic rocprofv3 measured are reported beside it.
def roofline(bytes_per_launch, launch_s, kernel, traffic_name, launch=None, kernel_bytes=None):
    ach = bytes_per_launch / launch_s / 1e9
    r = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(traffic_name, launch),
         "kernel": kernel,
         "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": round(launch_s * 1e6, 2)}
    if kernel_bytes is not None:
        r["kernel_bytes_per_launch"] = kernel_bytes
        r["kernel_bytes_frac"] = round(kernel_bytes / launch_s / 1e9 / HBM_PEAK_GBS, 4)
    return r


def line(metric, value, unit, steps, warmup, ms, dtype, data, config, roof, cpu):
    return {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms / steps, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": data, "config": config, "roofline": roof, "cpu_baseline": cpu,
            "host": {"nproc": os.cpu_count(), "cpu": cpu_model()}}


def fp64_form_name():
    from phix_amd import capi
    return "exact (reference order)" if capi.lib().phx_ch_get_fp64_exact() else "fma, stage states and b (12 passes)"


# ------------------------------------------------------------ workloads ----
# CH RK4 launch kinds (phx_ch_timing): 0 one-stage, 1/2 two-stage (1,2)/(3,4),
# 3 2D whole step. Algorithmic passes per launch (accounting above):
# one-stage 16/4, two-stage 16/2, whole step 16.
CH_ALG = {0: 16 / 4, 1: 16 / 2, 2: 16 / 2, 3: 16}


def ch_rk4_run(a, dtype_name, dim):
    """Times the fused CH RK4 path of one precision: the 256^3 zero-flux
    problem (dim 3) or the 1024^2 periodic one (dim 2); returns the workload
    fields of the JSON line."""
    import torch
    from phix_amd import capi, init
    dev = torch.device("cuda:0")
    dtype = capi.F32 if dtype_name == "f32" else capi.F64
    tdt, esz = (torch.float32, 4) if dtype == capi.F32 else (torch.float64, 8)
    n = (256,) * 3 if dim == 3 else (1024, 1024)
    bc = capi.NOFLUX if dim == 3 else capi.PERIODIC
    g = capi.make_grid(n, bc=bc)
    p = capi.CHParams(1.0, 0.0, 1.0, 1.0, 1.0, 0.005)
    c = torch.from_numpy(init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)).to(tdt).to(dev)
    work = torch.empty(capi.lib().phx_ch_rk4_workspace_bytes(C.byref(g), dtype) // esz, dtype=tdt, device=dev)
    stream = torch.cuda.current_stream()
    sh = capi.stream_handle(stream)
    zc, tune_ms = C.c_int(0), C.c_double(0)
    if a.autotune:
        scratch = c.clone()
        capi.call("phx_ch_rk4_autotune", C.byref(g), dtype, capi.ptr(scratch), capi.ptr(work), C.byref(p), 10,
                  C.byref(zc), C.byref(tune_ms), sh)
        del scratch
    if a.tile is not None:
        capi.call("phx_ch_rk4_set_tile", C.byref(g), dtype, a.tile)
    if a.zc is not None and dim == 3:
        capi.call("phx_ch_rk4_set_zchunk", C.byref(g), dtype, a.zc)
    tile = C.c_int(0)
    capi.call("phx_ch_rk4_get_launch", C.byref(g), dtype, C.byref(zc), C.byref(tile))

    def run(k):
        capi.call("phx_ch_rk4_steps", C.byref(g), dtype, capi.ptr(c), capi.ptr(work), C.byref(p), k, sh)

    run(a.warmup)
    ms, reps = timed_reps(run, a.steps, a.reps, stream)
    kinds = launch_shares(kernel_times(run, a.steps), ms)
    cells = math.prod(n)
    if dim == 3:
        # the kernels' own passes: one-stage 13/4 (FP32 stage-state form),
        # 12/4 (default FP64 form: stage 2 writes b = u2 + 2 u3 - c_n, stage 4
        # reads u4 and b) or 17/4 (FP64 exact form: the reference's
        # k1+2k2+2k3+k4 order re-reads c_n in stage 4); two-stage (1,2): read
        # c_n, write the stage-3 state and acc/b/u2; (3,4): read the stage-3
        # state, c_n and acc/b/u2, write c_{n+1}
        exact = dtype == capi.F64 and capi.lib().phx_ch_get_fp64_exact()
        own = {0: (17 if exact else (12 if dtype == capi.F64 else 13)) / 4, 1: 3, 2: 4}
        names = {0: "k_ch_stage_3dz" if tile.value >= 9 else "k_ch_stage_3d", 1: "k_ch_pair_3d (stages 1+2)",
                 2: "k_ch_pair_3d (stages 3+4)"}
        tags = {0: "ch3d_rk4", 1: "ch3d_pair12", 2: "ch3d_pair34"}
        launch = TILE_NAMES[tile.value]
        traffic_tag = f"{tags[max(kinds, key=lambda k: kinds[k][0] * kinds[k][1])]}_{dtype_name}_256"
    else:
        own = {0: 4, 3: 2}
        names = {0: "k_ch_stage_2d", 3: "k_ch_step_2d"}
        launch = TILE2_NAMES[tile.value]
        dk = max(kinds, key=lambda k: kinds[k][0] * kinds[k][1])
        traffic_tag = f"{'ch2d_rk4' if dk == 0 else 'ch2d_step'}_{dtype_name}_1024"
    dom = max(kinds, key=lambda k: kinds[k][0] * kinds[k][1])
    roof = roofline(CH_ALG[dom] * esz * cells, kinds[dom][1] / 1e3, names[dom], traffic_tag, launch,
                    own[dom] * esz * cells)
    roof["avg_launch_method"] = ("launch durations inside the median timed window: HIP-event durations of the "
                                 "same steps run again, scaled to the window (launches x average <= window)")
    roof["launches"] = {names[k]: {"count": v[0], "avg_us": round(v[1] * 1e3, 2),
                                   "event_avg_us": round(v[2] * 1e3, 2),
                                   "alg_bytes": CH_ALG[k] * esz * cells, "kernel_bytes": own[k] * esz * cells}
                        for k, v in kinds.items()}
    if dtype == capi.F64 and (dim == 3 or capi.lib().phx_ch_get_fp64_exact()):
        form = fp64_form_name()
    else:
        form = "fma, stage states" if dim == 3 else "fma, accumulator"
    return {"value": cells * a.steps / (ms / 1e3), "ms": ms, "reps_ms": reps, "roofline": roof,
            "launch": launch, "zchunk": zc.value, "form": form, "grid": g, "params": p}


def ch_cpu_baseline(g, p, n, steps):
    """The FP64 oracle (a restatement of the reference's per-term CPU path) on
    one host thread, then on every host core, over `steps` RK4 steps."""
    import numpy as np
    from phix_amd import init
    from oracle import oracle as O
    cells = math.prod(n)
    O.set_threads(1)   # the reference's CPU path is serial
    cc = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    w = np.zeros((6,) + cc.shape)
    t0 = time.perf_counter()
    O.ch_rk4_steps(g, p, cc, steps, w)
    dt = time.perf_counter() - t0
    shape = "x".join(map(str, n))
    cpu = {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
           "sample": f"{steps} RK4 steps of the same {shape} problem, FP64, reference-structured per-term passes "
                     f"(oracle/phx_oracle.c orc_ch_rk4_steps), {dt:.1f} s on {cpu_model()}"}
    # the same passes on every host core (OpenMP over k, j; same results)
    nt = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    cc = init.to_stored(init.ch_field(n, 2, 0.4, 0.02), g)
    t0 = time.perf_counter()
    O.ch_rk4_steps(g, p, cc, steps, w)
    dta = time.perf_counter() - t0
    O.set_threads(1)
    cpu["all_cores"] = {"value": cells * steps / dta, "cores": nt, "sample": f"the same {steps} steps, {dta:.1f} s"}
    return cpu


def bench_ch_rk4(a, with_cpu, dim):
    """The line of one precision (a.dtype) and, unless --no-companion, the
    same workload in the other precision as a companion object."""
    main = ch_rk4_run(a, a.dtype, dim)
    n = (256,) * 3 if dim == 3 else (1024, 1024)
    comp = None
    if not a.no_companion:
        other = "f32" if a.dtype == "f64" else "f64"
        r = ch_rk4_run(a, other, dim)
        comp = {"dtype": other, "value": round(r["value"], 1), "unit": "cell-updates/s",
                "ms_per_step": round(r["ms"] / a.steps, 5), "reps_ms": r["reps_ms"], "form": r["form"],
                "launch": r["launch"], "roofline": r["roofline"]}
        if dim == 3:
            comp["zchunk"] = r["zchunk"]
    cpu = ch_cpu_baseline(main["grid"], main["params"], n, 10 if dim == 3 else 20) if with_cpu else None
    if dim == 3:
        cfg = {"workload": "ch3d_rk4_256cube_noflux", "grid": list(n), "bc": "noflux", "scheme": "rk4", "dt": 0.005,
               "kappa": 1.0, "M": 1.0, "f": "c^2(1-c)^2", "bulk": a.dtype, "form": main["form"],
               "parallelism": "single-gpu", "zchunk": main["zchunk"], "tile": main["launch"]}
        metric = "cell-updates/s, Cahn-Hilliard RK4 (fused stencil, %HBM roofline in 'roofline')"
    else:
        cfg = {"workload": "ch2d_rk4_1024sq_periodic", "grid": list(n), "bc": "periodic", "scheme": "rk4",
               "dt": 0.005, "kappa": 1.0, "M": 1.0, "f": "c^2(1-c)^2", "bulk": a.dtype, "form": main["form"],
               "parallelism": "single-gpu", "tile": main["launch"]}
        metric = "cell-updates/s, Cahn-Hilliard RK4 2D (fused stencil)"
    out = line(metric, main["value"], "cell-updates/s", a.steps, a.warmup, main["ms"], a.dtype,
               "synthetic (c0 = 0.4 + U(-0.02, 0.02), SplitMix64 seed 2)", cfg, main["roofline"], cpu)
    out["timing"] = {"reps": a.reps, "steps_per_rep": a.steps, "reps_ms": main["reps_ms"],
                     "value_from": "median window"}
    if comp is not None:
        out["companion"] = comp
    return out


def bench_ch3d(a, with_cpu):
    return bench_ch_rk4(a, with_cpu, 3)


def bench_ch2d(a, with_cpu):
    return bench_ch_rk4(a, with_cpu, 2)


def bench_ac(a, with_cpu):
    import torch
    from phix_amd import capi, init
    NX, NF = 1024, 8
    dev = torch.device("cuda:0")
    g = capi.make_grid((NX, NX), bc=capi.PERIODIC)
    p = capi.ACParams(1.0, 2.0, 1.5, 0.01)
    x = init.ac_fields((NX, NX), NF, 3)
    etas = [torch.from_numpy(init.to_stored(x[f], g)).float().to(dev) for f in range(NF)]
    arr = (C.c_void_p * NF)(*[capi.ptr(t) for t in etas])
    work = torch.empty(capi.lib().phx_ac_rk4_workspace_bytes(C.byref(g), capi.F32, NF) // 4, dtype=torch.float32,
                       device=dev)
    stream = torch.cuda.current_stream()
    sh = capi.stream_handle(stream)

    def run(n):
        capi.call("phx_ac_rk4_steps", C.byref(g), capi.F32, NF, arr, capi.ptr(work), C.byref(p), n, sh)

    run(a.warmup)
    ms = timed(lambda: run(a.steps), stream)
    kinds = kernel_times(run, a.steps)
    pts = NX * NX
    # algorithmic passes per launch and field (accounting above): one-stage
    # 16/4, two-stage 16/2. The two-stage kernels' own passes: (1,2) read c_n,
    # write acc and the stage-3 state; (3,4) read the stage-3 state, c_n and
    # acc, write c_{n+1}
    alg = {4: 16 / 4, 5: 16 / 2, 6: 16 / 2}
    own = {4: 16 / 4, 5: 3, 6: 4}
    names = {4: "k_ac_stage_2d", 5: "k_ac_pair_2d (stages 1+2)", 6: "k_ac_pair_2d (stages 3+4)"}
    launch = {4: "one-stage", 5: "two-stage (1,2)", 6: "two-stage (3,4)"}
    dom = max(kinds, key=lambda k: kinds[k][1])
    roof = roofline(alg[dom] * NF * 4 * pts, kinds[dom][1] / 1e3 / kinds[dom][0], names[dom], "ac8_rk4_f32_1024",
                    launch[dom], own[dom] * NF * 4 * pts)
    roof["launches"] = {names[k]: {"count": v[0], "avg_us": round(v[1] * 1e3 / v[0], 2),
                                   "alg_bytes": alg[k] * NF * 4 * pts, "kernel_bytes": own[k] * NF * 4 * pts}
                        for k, v in kinds.items()}
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        O.set_threads(1)   # the reference's CPU path is serial
        steps = 1
        es = [init.to_stored(x[f], g) for f in range(NF)]
        t0 = time.perf_counter()
        O.ac_rk4_steps(g, p, es, steps)
        dt = time.perf_counter() - t0
        cpu = {"value": pts * steps / dt, "unit": "point-updates/s", "cores": 1, "kind": "port",
               "sample": f"{steps} RK4 step of the same 1024^2 x 8-field problem, FP64 oracle, {dt:.1f} s on "
                         f"{cpu_model()}"}
    cfg = {"workload": "ac8_rk4_1024sq_periodic", "grid": [NX, NX], "fields": NF, "gamma": 1.5, "kappa": 2.0,
           "L": 1.0, "dt": 0.01, "scheme": "rk4", "bulk": "f32", "parallelism": "single-gpu"}
    return line("grid-point-updates/s (8 fields), Allen-Cahn RK4 (fused multi-field stencil)",
                pts * a.steps / (ms / 1e3), "point-updates/s", a.steps, a.warmup, ms, "f32",
                "synthetic (one-hot on floor(8u) + U(-0.001, 0.001), SplitMix64 seed 3)", cfg, roof, cpu)


def bench_ens(a, with_cpu):
    import torch
    import numpy as np
    from phix_amd import capi, init
    dev = torch.device("cuda:0")
    N, kappa, cbar = init.ensemble_params(512, 4)
    n = len(N)
    h = C.c_void_p()
    capi.call("phx_ensemble_create", n, (C.c_int * n)(*[int(v) for v in N]), (C.c_double * n)(*kappa), None,
              C.byref(h))
    cells = capi.lib().phx_ensemble_cells(h)
    x0 = np.concatenate([init.uniform(init.substream(4, 10 + r), int(N[r]) ** 2, cbar[r] - 0.01, cbar[r] + 0.01)
                         for r in range(n)])
    c = torch.from_numpy(x0).float().to(dev)
    work = torch.empty(3 * cells, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    sh = capi.stream_handle(stream)
    scratch = c.clone()
    best = C.c_double(0)
    capi.call("phx_ensemble_autotune", h, capi.ptr(scratch), capi.ptr(work), 0.01, 5, C.byref(best), sh)
    del scratch
    sp, gr, tx, ty = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    capi.call("phx_ensemble_get_tiling", h, C.byref(sp), C.byref(gr), C.byref(tx), C.byref(ty))
    per_graph = 50
    steps = max(per_graph, (a.steps // per_graph) * per_graph)
    capi.call("phx_ensemble_graph_steps", h, capi.ptr(c), capi.ptr(work), 0.01, per_graph, 1, sh)
    ms = timed(lambda: capi.call("phx_ensemble_graph_steps", h, capi.ptr(c), capi.ptr(work), 0.01, per_graph,
                                 steps // per_graph, sh), stream)
    two = capi.lib().phx_ensemble_get_two_stage(h)
    # per-launch durations: the same steps again, eager, each launch bracketed
    # by HIP events (outside the timed region); field passes per launch as
    # for the Allen-Cahn kernels
    kinds = kernel_times(lambda k: capi.call("phx_ensemble_rk4_steps", h, capi.ptr(c), capi.ptr(work), 0.01, k, sh),
                         steps)
    alg = {7: 16 / 4, 8: 16 / 2, 9: 16 / 2}
    own = {7: 16 / 4, 8: 3, 9: 4}
    names = {7: "k_ens_stage", 8: "k_ens_pair (stages 1+2)", 9: "k_ens_pair (stages 3+4)"}
    launch = {7: "one-stage", 8: "two-stage (1,2)", 9: "two-stage (3,4)"}
    dom = max(kinds, key=lambda k: kinds[k][1])
    roof = roofline(alg[dom] * 4 * cells, kinds[dom][1] / 1e3 / kinds[dom][0], names[dom], "ens512_rk4_f32",
                    launch[dom], own[dom] * 4 * cells)
    roof["launches"] = {names[k]: {"count": v[0], "avg_us": round(v[1] * 1e3 / v[0], 2),
                                   "alg_bytes": alg[k] * 4 * cells, "kernel_bytes": own[k] * 4 * cells}
                        for k, v in kinds.items()}
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        O.set_threads(1)   # the reference's CPU path is serial
        csteps, runs = 20, 16
        done, off = 0, 0
        t0 = time.perf_counter()
        for r in range(runs):
            nn = int(N[r])
            g = capi.make_grid((nn, nn), bc=capi.PERIODIC)
            cc = init.to_stored(x0[off:off + nn * nn].reshape(1, nn, nn), g)
            off += nn * nn
            O.ch_rk4_steps(g, capi.CHParams(1, 0, 1, float(kappa[r]), 1, 0.01), cc, csteps)
            done += nn * nn * csteps
        dt = time.perf_counter() - t0
        cpu = {"value": done / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"{csteps} RK4 steps of the first {runs} runs, FP64 oracle, {dt:.1f} s on {cpu_model()}"}
    capi.lib().phx_ensemble_destroy(h)
    cfg = {"workload": "ens512_ch_rk4_periodic", "runs": n, "cells": int(cells), "sides": "64+32*floor(15u)",
           "dt": 0.01, "scheme": "rk4", "bulk": "f32", "tile": [tx.value, ty.value],
           "kernel": ("two-stage" if two == 1 else "one-stage") + f", persistent grid {gr.value}",
           "hipgraph_steps": per_graph, "parallelism": "single-gpu"}
    return line("cell-updates/s, batched uneven-size CH RK4 ensemble", cells * steps / (ms / 1e3),
                "cell-updates/s", steps, per_graph, ms, "f32",
                "synthetic (per-run kappa ~ U(0.25,1), cbar ~ U(0.3,0.7) +- 0.01, SplitMix64 seed 4)", cfg, roof, cpu)


def bench_euler(a, with_cpu):
    import torch
    import numpy as np
    from phix_amd import capi, init
    NX = 128
    dev = torch.device("cuda:0")
    g = capi.make_grid((NX, NX), bc=capi.PERIODIC)
    p = capi.CHParams(1.0, 0.0, 1.0, 0.5, 1.0, 0.01)
    c0 = init.to_stored(init.ch_field((NX, NX), 1, 0.5, 0.01), g)
    c = torch.from_numpy(c0).to(dev)
    mu = torch.zeros_like(c)
    stream = torch.cuda.current_stream()
    sh = capi.stream_handle(stream)
    steps = 1000

    def run(n):
        capi.call("phx_ch_euler_steps", C.byref(g), capi.F64, capi.ptr(c), capi.ptr(mu), C.byref(p), n, sh)

    run(10)
    ms = timed(lambda: run(steps), stream)
    cells = NX * NX
    # steps 1..996 of the call run as 249 launches of four steps each (the
    # whole-step kernel's Euler form), then three one-step launches and the
    # last step's k_euler_mu + k_euler_c: a launch is credited with four
    # steps of read c, write c' and given four steps' share of the window
    roof = roofline(4 * 2 * 8 * cells, 4 * (ms / 1e3) / steps, "k_ch_step_2d (Euler form, four steps per launch)",
                    "ch2d_euler_f64_128")
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        O.set_threads(1)   # the reference's CPU path is serial
        cc = c0.copy()
        t0 = time.perf_counter()
        O.ch_euler_steps(g, p, cc, np.zeros_like(cc), steps)
        dt = time.perf_counter() - t0
        cpu = {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"the full {steps}-step run, FP64 oracle, {dt:.1f} s on {cpu_model()}"}
    cfg = {"workload": "ch2d_euler_128sq_periodic", "grid": [NX, NX], "bc": "periodic", "scheme": "euler multi-step",
           "dt": 0.01, "kappa": 0.5, "M": 1.0, "bulk": "f64", "parallelism": "single-gpu",
           "launch": "four steps per launch, 32x16 cells per 512-thread workgroup",
           "note": "135 KB working set is cache resident: latency bound, not HBM bound"}
    return line("cell-updates/s, Cahn-Hilliard multi-step Euler (config-1 app)", cells * steps / (ms / 1e3),
                "cell-updates/s", steps, 10, ms, "f64", "synthetic (c0 = 0.5 + U(-0.01, 0.01), SplitMix64 seed 1)",
                cfg, roof, cpu)


def bench_gfa(a, with_cpu):
    """The reference's FeB_ex1 case of the GFA_2D program (Fe-B glass
    formation): 600^2 cells of 0.3 mm, zero-flux faces, dt = 1 us, its
    constants, a B-rich crystal nucleus of radius 2 cells (c = 0.33 +
    U(-0.01, 0.01), phi = 1, eta = 0) in an amorphous matrix (c = 0.25,
    phi = 0, eta = 1), as its gen_initial_field.py lays out. One fused launch
    per step (phx_gfa_euler_steps)."""
    import torch
    import numpy as np
    from phix_amd import capi
    NX = 600
    DX = 3.0e-4
    k = {"M_eta": 1.0, "D_Am": 3.0e-12, "eps_sq": 2.86e-3, "beta_sq": 1.0e-4, "w_phi": 441.0, "w_eta": 10.0,
         "w_ex": 100.0, "T": 1000.0, "T_g": 710.0, "alpha": 0.45, "V_m_L": 2.64e-5, "V_m_S": 2.44e-5}
    dev = torch.device("cuda:0")
    g = capi.make_grid((NX, NX), d=(DX, DX), bc=capi.NOFLUX)
    p = capi.GfaParams(dt=1.0e-6, **k)
    yy, xx = np.mgrid[0:NX, 0:NX]
    inside = (xx + 0.5 - NX / 2) ** 2 + (yy + 0.5 - NX / 2) ** 2 <= 2.0 ** 2
    noise = np.random.default_rng(42).uniform(-0.01, 0.01, (NX, NX))
    phys = [np.where(inside, 0.33 + noise, 0.25), np.where(inside, 1.0, 0.0), np.where(inside, 0.0, 1.0)]
    fields = []
    for f in phys:
        a3 = np.zeros(capi.stored_shape(g))
        capi.physical(a3, g)[0] = f
        fields.append(torch.from_numpy(a3).to(dev))
    work = torch.zeros(capi.lib().phx_gfa_workspace_bytes(C.byref(g)) // 8, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = capi.stream_handle(stream)
    steps = 8 * a.steps   # even (the state ends in place); default 4000, a launch is ~12 us

    def run(n):
        capi.call("phx_gfa_euler_steps", C.byref(g), *(capi.ptr(f) for f in fields), capi.ptr(work), C.byref(p), n,
                  sh)

    run(a.warmup)
    ms = timed(lambda: run(steps), stream)
    cells = NX * NX
    # The timed region is one k_gfa_step launch per step on the timed stream
    # (an even step count: no final copy), so its average is the launch
    # duration; events around each ~12 us launch would add ~2 us.
    # Per cell-step: read c, phi, eta and write them, FP64; mu and D stay on chip.
    roof = roofline(48 * cells, (ms / 1e3) / steps, "k_gfa_step", "gfa2d_feb_f64_600", kernel_bytes=48 * cells)
    cpu = None
    if with_cpu:
        from oracle import gfa as G
        sample = 40
        t0 = time.perf_counter()
        G.steps(G.Model(k), *phys, DX, DX, ("NoFlux", "NoFlux"), 1.0e-6, sample)
        dt = time.perf_counter() - t0
        cpu = {"value": cells * sample / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"{sample} steps of the same 600^2 problem, FP64 numpy oracle (oracle/gfa.py), {dt:.1f} s "
                         f"on {cpu_model()}"}
    cfg = {"workload": "gfa2d_feb_ex1_600sq_noflux", "grid": [NX, NX], "dx": DX, "bc": "noflux",
           "scheme": "euler, six multi-steps per step", "dt": 1.0e-6, "T": 1000.0, "bulk": "f64",
           "kernel": "k_gfa_step 32x16", "parallelism": "single-gpu",
           "note": "three 2.9 MB fields stay cache resident between steps"}
    return line("cell-updates/s, Fe-B glass-formation (GFA_2D) Euler step", cells * steps / (ms / 1e3),
                "cell-updates/s", steps, a.warmup, ms, "f64",
                "synthetic (FeB_ex1 layout: radius-2 nucleus, c noise U(-0.01, 0.01) from numpy default_rng(42))",
                cfg, roof, cpu)


def bench_surrogate(a, with_cpu):
    """Config 5: the surrogate-plus-solver form (parareal after a plain-solver
    lead phase; coarse = the BF16 MFMA network, fine = the batched FP32 RK4
    solver) against the plain FP32 GPU solver, 256^2 periodic CH to t_end = 200
    (20,000 RK4 steps), over a fixed set of initial conditions, each final
    field's relative L2 against the FP64 GPU solver. `value` is the median
    speedup in time to solution; `roofline` is the fused network forward's."""
    import torch
    import numpy as np
    from phix_amd import capi, surrogate as S
    dev = torch.device("cuda:0")
    params = np.load(S.WEIGHTS_LATE)["params"]
    ics = S.EVAL_ICS[: a.ics]
    ev = S.evaluate_parareal(dev, params, ics=ics, reps=3)
    # the fused forward: a 190-jump rollout (one launch per jump, one
    # projection at the end) timed with events on the launch stream
    net = S.Surrogate(dev, 1, params=params)
    x = torch.from_numpy(S.config5_initial(ics[0])).float().to(dev).reshape(1, 256, 256).contiguous()
    bufs = [torch.empty_like(x) for _ in range(2)]
    net.forward(x)
    jumps = 190
    stream = torch.cuda.current_stream()
    net.rollout(x, jumps, bufs)
    ms = min(timed(lambda: net.rollout(x, jumps, bufs), stream) for _ in range(3))
    per_jump = ms / 1e3 / jumps
    px = 256 * 256
    flops = px * (2 * 9 * 32 + 4 * 2 * 9 * 32 * 32 + 2 * 9 * 32)   # the network's multiply-adds x 2 per jump
    ach = flops / per_jump / 1e12
    roof = {"bound": "mfma", "achieved": round(ach, 1), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(ach / 2500.0, 4), "traffic": None, "kernel": "k_sur_fused_fwd",
            "alg_flops_per_launch": flops, "avg_launch_us": round(per_jump * 1e6, 2),
            "note": "network FLOPs of one batch-1 256^2 jump (BF16 MFMA dense peak); the 16x16 tiles recompute "
                    "a 6-pixel ring, 1.74x these FLOPs in layers 2-5; average over a 190-jump rollout "
                    "(190 fused launches + 1 projection)"}
    cpu = None
    if with_cpu:
        from phix_amd import init
        from oracle import oracle as O
        O.set_threads(1)
        g = capi.make_grid((256, 256), bc=capi.PERIODIC)
        pp = capi.CHParams(1.0, 0.0, 1.0, 0.5, 1.0, 0.01)
        cc = init.to_stored(S.config5_initial(ics[0]), g)
        w = np.zeros((6,) + cc.shape)
        steps = 2000
        t0 = time.perf_counter()
        O.ch_rk4_steps(g, pp, cc, steps, w)
        dt = time.perf_counter() - t0
        cpu = {"value": round(dt * 20000 / steps, 2), "unit": "s to t_end (plain solver, 20,000 RK4 steps)",
               "cores": 1, "kind": "port",
               "sample": f"{steps} RK4 steps of the same 256^2 problem, FP64 oracle (orc_ch_rk4_steps), "
                         f"{dt:.1f} s on {cpu_model()}, scaled to 20,000 steps"}
    pms = float(np.median([r["parareal_ms"] for r in ev["per_ic"]]))
    cfg = {"workload": "surrogate_ch256_tend200_periodic", "grid": [256, 256], "bc": "periodic", "kappa": 0.5, "M": 1.0,
           "dt": 0.01, "t_end_steps": 20000, "jump_steps": 100, "lead_steps": ev["lead_steps"], "slices": ev["slices"],
           "slice_steps": ev["slice_steps"], "iterations": 1, "weights": S.WEIGHTS_LATE.name, "ics": list(ics),
           "plain_solver": "fused FP32 RK4, autotuned 2D launch", "parallelism": "single-gpu"}
    out = line("speedup in time to t_end of the surrogate-plus-solver form over the plain GPU solver (256^2 CH, "
               "median over initial conditions; relative L2 vs FP64 per IC in 'surrogate')",
               ev["speedup_median"], "x", 1, 1, pms, "bf16+f32",
               "synthetic (c0 = 0.5 + U(-0.05, 0.05), SplitMix64 substream (1234, ic))", cfg, roof, cpu)
    out["value"] = round(ev["speedup_median"], 3)
    out["surrogate"] = ev
    return out


def bench_copy(a, with_cpu):
    """Plain device copy of 2 GiB (the STREAM-copy bandwidth of this box, and the
    known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    n = (2 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    dst.copy_(src)
    reps = 10
    ms = timed(lambda: [dst.copy_(src) for _ in range(reps)], stream)
    per = ms / 1e3 / reps
    roof = roofline(2.0 * 4 * n, per, "copy (torch)", "copy_2gib")
    cfg = {"workload": "copy_2gib", "bytes_read": 4 * n, "bytes_written": 4 * n, "parallelism": "single-gpu"}
    return line("GB/s, device-to-device copy (read + write bytes / time)", 2.0 * 4 * n / per / 1e9, "GB/s", reps, 1,
                ms, "f32", "constant buffer", cfg, roof, None)


WORKLOADS = {"copycal": bench_copy, "ch3d": bench_ch3d, "ch2d1024": bench_ch2d, "ac1024": bench_ac, "ens512": bench_ens,
             "euler128": bench_euler, "gfa600": bench_gfa, "surrogate": bench_surrogate}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f64",
                    help="CH RK4 precision of the line (the other one is reported as 'companion')")
    ap.add_argument("--reps", type=int, default=5, help="timed windows of --steps steps; the median is reported")
    ap.add_argument("--no-companion", action="store_true", help="skip the other precision's companion run")
    ap.add_argument("--autotune", action="store_true", help="CH RK4: run the launch autotuner first")
    ap.add_argument("--fp64-exact", action="store_true",
                    help="CH RK4 FP64 in the reference's operation order (bit-equal to the oracle)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="ch3d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", type=int, default=None, help="CH RK4 launch tile index (after any autotuning)")
    ap.add_argument("--zc", type=int, default=None, help="3D stage z-chunk (after any autotuning)")
    ap.add_argument("--ics", type=int, default=10, help="surrogate: initial conditions evaluated (up to 10)")
    a = ap.parse_args(argv)
    a.reps = max(1, a.reps)
    if a.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for --gpus {a.gpus}); "
              "the stencil path shards by slab with halo exchange but that is not built here.", file=sys.stderr)
        return 2
    if a.fp64_exact:
        from phix_amd import capi
        capi.call("phx_ch_set_fp64_exact", 1)
    out = WORKLOADS[a.workload](a, not a.no_cpu_baseline)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
