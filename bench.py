"""Throughput of the fluid step on one MI355X.

Workload (default): the 256^3 f32 smoke plume, open top, buoyancy 1, sphere
source, 5 advect+project steps at dt = 0.5/256 with MG-PCG (4 levels, 2+2
red-black sweeps) to 1e-6 relative L-inf residual. Inputs (seeded
band-limited velocity, zero density) are generated on the device and resident
in HBM before the timed region. A step is one call of pano_fluid_step (it
returns when the step is complete).

Prints one JSON line: cell-updates/s over the timed steps, step latency
p50/p99, the roofline of the dominant kernel (algorithmic bytes per launch /
average launch duration measured with HIP events on the library's stream
during the timed steps), the accuracy of the projected velocity (FP64 on the
host from the step's outputs), and the CPU oracle timed on a bounded sample
of the same workload on this host (one thread and all cores), with its FP64
error check. Unless --no-extra is given, the same line also carries the other
workloads, each measured in this run: the 512^3 mixed-precision Poisson solve
("mixed_poisson512"), the 64-grid batch ("batch64_2d") and the served
projection mix ("serve", 60 s of arrivals).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel smooth_post_l0] [--no-cpu] [--no-extra]
    python bench.py --workload {poisson512_mixed,batch64_2d,serve,stream_copy}   (one workload alone)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="smooth_post_l0", help="kernel whose roofline is reported")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--dims", default=None)
    ap.add_argument("--workload", default="smoke3d_256",
                    choices=["smoke3d_256", "poisson512_mixed", "batch64_2d", "serve", "stream_copy"])
    ap.add_argument("--duration", type=float, default=None, help="serve: seconds of arrivals (default 60)")
    ap.add_argument("--cpu-check", action="store_true",
                    help="poisson512_mixed: also solve in pure FP64 with the CPU oracle (minutes at 512^3)")
    ap.add_argument("--no-extra", action="store_true",
                    help="default run: leave out the mixed, batch and serving workloads")
    args = ap.parse_args()
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 5
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the fluid step runs on a single MI355X); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "poisson512_mixed":
        return print(json.dumps(bench_mixed(args)), flush=True)
    if args.workload == "batch64_2d":
        return print(json.dumps(bench_batch(args)), flush=True)
    if args.workload == "stream_copy":
        return bench_stream_copy(args)
    if args.workload == "serve":
        return print(json.dumps(bench_serve(args)), flush=True)

    import panopaea_amd as pa
    from panopaea_amd import _lib
    from panopaea_amd import workloads as W

    wl = W.SMOKE3D_256
    dims = tuple(int(v) for v in (args.dims or "256,256,256").split(","))
    g = W.make_grid(wl, dims)
    cfg = W.step_config(wl, dims)
    n = g.num_elem_2()
    ctx = g._ctx()
    L = _lib.lib()

    rho = [g.new_simplex_2(), g.new_simplex_2()]
    vel = [g.new_simplex_1(), g.new_simplex_1()]
    press = g.new_simplex_2()
    t0 = time.time()
    pa.gen_bandlimited_faces_device(vel[0], wl["seed"], wl["n_modes"], wl["k_max"])
    print(f"bench: {wl['name']} dims={dims} generated inputs in {time.time() - t0:.2f}s", file=sys.stderr,
          flush=True)
    init_vel = vel[0].download() if not args.no_cpu else None
    params = cfg.params()

    cur = 0

    def step():
        nonlocal cur
        info = pa.fluid_step(g, cfg, rho[cur], vel[cur], rho[1 - cur], vel[1 - cur], press, params=params)
        cur = 1 - cur
        return info

    for k in range(args.warmup):
        info = step()
        print(f"bench: warmup {k} iterations={info.iterations} converged={info.converged}", file=sys.stderr,
              flush=True)

    _lib.check(L.pano_timer_start(ctx.handle, args.kernel.encode()))
    ctx.synchronize()
    lat, iters, infos = [], [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        a = time.perf_counter()
        info = step()
        lat.append(time.perf_counter() - a)
        iters.append(info.iterations)
        infos.append(info.as_dict())
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    tot_ms, launches, bpl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    _lib.check(L.pano_timer_stop(ctx.handle, ctypes.byref(tot_ms), ctypes.byref(launches), ctypes.byref(bpl)))
    for k, i in enumerate(infos):
        print(f"bench: step {k} {lat[k] * 1e3:.2f} ms iterations={i['iterations']} converged={i['converged']} "
              f"r_linf/b_linf={i['r_linf'] / max(i['b_linf'], 1e-300):.2e}", file=sys.stderr, flush=True)

    value = n * args.steps / elapsed
    roof = None
    if launches.value > 0:
        avg_s = tot_ms.value / launches.value / 1e3
        achieved = bpl.value / avg_s / 1e9
        traffic = profiled_traffic(args.kernel, dims)
        roof = {"bound": "hbm", "kernel": args.kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "launches": launches.value, "avg_us": round(avg_s * 1e6, 2),
                "bytes_per_launch": bpl.value}
        if traffic:
            # the bytes the launch really moves (counter-calibrated) over the same time
            roof["traffic_gbs"] = round(traffic / avg_s / 1e9, 1)
            roof["traffic_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)

    # the whole step against the same unit rule: model bytes of the reference
    # method (phase-separated passes) over the measured step time
    # accuracy of the last timed step, from its outputs (FP64 on the host)
    from panopaea_amd import diagnostics as D
    last = infos[-1]
    rep = D.projection_report(g, vel[cur].download(), press.download(), cfg.dt, last["b_linf"])
    accuracy = {
        "recursive_rel_residual_linf": last["r_linf"] / last["b_linf"],
        "true_rel_residual_linf": rep["div_rel_linf"],
        "true_residual_l2": rep["div_l2"],
        "rounding_bound_rel": rep["rounding_bound_rel"],
        "note": "true = ||div u_new||_inf / ||b||_inf in FP64 from the step's f32 outputs: the solve's b - A p "
                "up to the f32 rounding of the projected faces (rounding_bound_rel); last timed step",
    }
    print(f"bench: accuracy {accuracy}", file=sys.stderr, flush=True)

    k_mean = float(np.mean(iters))
    b_step = step_model_bytes(n, k_mean)
    step_gbs = b_step * args.steps / elapsed / 1e9
    step_roof = {"bound": "hbm", "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(step_gbs / HBM_PEAK_GBS, 4), "model_bytes_per_step": b_step,
                 "cg_iterations_mean": k_mean,
                 "model": "4N(8 + 4 + 7 + 18.54 + 29.54 K) bytes: advect 8, divergence 4, gradient 7, V-cycle "
                          "18.54 and per CG iteration 11 + 18.54 f32 field passes"}

    cpu = None
    if not args.no_cpu:
        cpu, acc_fp64 = cpu_baseline(wl, dims, cfg, init_vel)
        accuracy["vs_fp64_oracle"] = acc_fp64
    del rho, vel, press

    out = {
        "metric": "cell-updates/s 256^3 advect+project @1e-6 resid, plain FP32; p50/p99 step latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded band-limited velocity, sphere density source)",
        "config": {"workload": wl["name"] if dims == tuple(wl["dims"]) else f"smoke3d_{dims}", "dims": list(dims),
                   "dt": cfg.dt, "precond": "multigrid 4 levels, 2+2 red-black GS", "tol": "1e-6 rel L-inf",
                   "bc": "closed walls, open top", "buoyancy": 1.0,
                   "cg_iterations_per_step": iters},
        "roofline": roof,
        "step_roofline": step_roof,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        args.steps_given, args.warmup = False, 2
        out["mixed_poisson512"] = extra(bench_mixed, args, reps=3)
        out["batch64_2d"] = extra(bench_batch, args)
        out["serve"] = extra(bench_serve, args)
    print(json.dumps(out), flush=True)


_COMMON = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "warmup")


def extra(fn, args, **kw):
    """Another workload measured in the same run, as a key of the default
    line (its own line's fields, without the ones the line already has); an
    error is reported in place of the result."""
    try:
        r = fn(args, **kw)
        return {k: v for k, v in r.items() if k not in _COMMON}
    except Exception as e:  # reported, not hidden: the key carries the error
        print(f"bench: {fn.__name__} failed: {e!r}", file=sys.stderr, flush=True)
        return {"error": repr(e)}


def step_model_bytes(n, k):
    """Algorithmic bytes of one 3D MG-PCG fluid step with k CG iterations, in
    f32 field passes of n cells: advect (read u, v, w, rho; write them) 8,
    divergence 4, gradient subtraction 7, the initial V-cycle 18.54 (16.25 per
    level over the three smoothed levels, geometric factor 1/8), and per
    iteration q = Ap with p.q 2, the x/r update with the norm 6, the search
    update 3 and a V-cycle 18.54."""
    return 4.0 * n * (8 + 4 + 7 + 18.54 + 29.54 * k)


def bench_mixed(args, reps=None):
    """512^3 Poisson: FP64 iterative refinement around the FP32 MG-PCG against
    the pure-FP64 MG-PCG (same FP64-verified residual) and the plain FP32
    MG-PCG (its attainable accuracy, measured in FP64). The plain FP32 form
    stalls near 1e-6 and cannot reach 1e-10, so the mixed form's speed is
    compared with the FP64 MG-PCG at the same verified residual. The CPU leg
    (unless --no-cpu) solves a bounded sample (the same generator at 128^3)
    in pure FP64 with the oracle and reports the mixed solve's error against
    it."""
    reps = reps or args.steps
    import panopaea_amd as pa
    from panopaea_amd import workloads as W
    wl = W.POISSON512_MIXED
    dims = tuple(int(v) for v in (args.dims or "512,512,512").split(","))
    n = int(np.prod(dims))
    g64 = pa.Grid3d(dims, np.float64, bc=wl["bc"])
    g32 = pa.Grid3d(dims, np.float32, bc=wl["bc"], ctx=g64._ctx())
    ctx = g64._ctx()
    faces = g64.new_simplex_1()
    pa.gen_uniform_faces_device(faces, wl["seed"])  # 24-bit values: the same in f32 and f64
    b64 = g64.new_simplex_2()
    pa.neg_divergence(b64, faces)
    faces.free()
    mg = pa.Multigrid(wl["mg_levels"], wl["mg_nu"])

    def timed(fn, reps):
        ts, out = [], None
        for _ in range(reps):
            ctx.synchronize()
            t = time.perf_counter()
            out = fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t)
        return out, ts

    # mixed precision
    x_mx = g64.new_simplex_2()
    for _ in range(args.warmup):
        pa.poisson_solve_mixed(x_mx, b64, dt=wl["dt"], preconditioner=mg, inner_tol=wl["inner_tol"],
                               outer_tol=wl["outer_tol"])
    info_mx, t_mx = timed(lambda: pa.poisson_solve_mixed(x_mx, b64, dt=wl["dt"], preconditioner=mg,
                                                         inner_tol=wl["inner_tol"], outer_tol=wl["outer_tol"]),
                          reps)
    print(f"bench: mixed {info_mx.as_dict()} times {[round(t * 1e3, 1) for t in t_mx]} ms", file=sys.stderr,
          flush=True)
    # pure FP64 MG-PCG to the same relative residual
    x64, r, a, s = (g64.new_simplex_2() for _ in range(4))
    op64 = pa.PoissonOperator(g64, wl["dt"])
    info_64, t_64 = timed(lambda: pa.precond_conjugate_gradient(mg, x64, b64, 1000, wl["outer_tol"], r, a, s, op64,
                                                                tol_kind="rel"), max(1, reps // 2))
    print(f"bench: fp64 pcg iterations={info_64.iterations} converged={info_64.converged} "
          f"times {[round(t * 1e3, 1) for t in t_64]} ms", file=sys.stderr, flush=True)
    # plain FP32 MG-PCG to 1e-6 (FP32 right-hand side)
    bh = b64.download()
    b32 = g32.new_simplex_2().upload(bh.astype(np.float32))
    x32, r32, a32, s32 = (g32.new_simplex_2() for _ in range(4))
    op32 = pa.PoissonOperator(g32, wl["dt"])
    info_32, t_32 = timed(lambda: pa.precond_conjugate_gradient(mg, x32, b32, 1000, 1e-6, r32, a32, s32, op32,
                                                                tol_kind="rel"), max(1, reps // 2))
    # accuracy, measured in FP64 on the device: ||b - A x||_inf / ||b||_inf and the error vs the FP64 solve
    def rel_res(xd):
        q = g64.new_simplex_2()
        op64(q, xd)
        return float(np.abs(bh - q.download()).max() / np.abs(bh).max())

    x32_64 = g64.new_simplex_2().upload(x32.download().astype(np.float64))
    xm, xf, xs = x_mx.download(), x64.download(), x32_64.download()
    ref = xf - xf.mean()
    err = lambda v: float(np.abs((v - v.mean()) - ref).max() / np.abs(ref).max())
    acc = {"mixed_rel_residual": rel_res(x_mx), "fp64_rel_residual": rel_res(x64), "fp32_rel_residual": rel_res(x32_64),
           "mixed_err_vs_fp64_gpu": err(xm), "fp32_err_vs_fp64_gpu": err(xs)}
    if args.cpu_check:
        from oracle import oracle as o
        P = o.params(dims, bc=wl["bc"], dt=wl["dt"], precond="mg", levels=wl["mg_levels"], tol=1e-12,
                     max_iterations=2000)
        t = time.perf_counter()
        xo, oinfo, _ = o.solve(P, bh)
        print(f"bench: cpu fp64 oracle solve {time.perf_counter() - t:.1f}s iterations={oinfo.iterations}",
              file=sys.stderr, flush=True)
        ref = xo - xo.mean()
        acc["mixed_err_vs_fp64_cpu"] = err(xm)
        acc["fp64_gpu_err_vs_fp64_cpu"] = err(xf)
    for f in (x_mx, x64, r, a, s, b32, x32, r32, a32, s32, x32_64, b64):
        f.free()
    del xm, xf, xs, bh, ref
    if not args.no_cpu:
        acc["sample_128_vs_fp64_cpu"] = mixed_cpu_check(pa, wl, mg)
    print(f"bench: accuracy {acc}", file=sys.stderr, flush=True)
    tm = float(np.median(t_mx))
    out = {
        "metric": "cells/s 512^3 Poisson to 1e-10 rel residual, FP64 iterative refinement around FP32 MG-PCG",
        "value": n / tm, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tm * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 outer", "data": "synthetic (seeded i.i.d. uniform[-1,1] faces)",
        "config": {"workload": wl["name"], "dims": list(dims), "inner_tol": wl["inner_tol"],
                   "outer_tol": wl["outer_tol"], "precond": "multigrid 4 levels, 2+2 red-black GS",
                   "outer_iterations": info_mx.outer_iterations, "inner_iterations": info_mx.inner_iterations},
        "fp64_pcg": {"ms": float(np.median(t_64)) * 1e3, "iterations": info_64.iterations},
        "fp32_pcg_1e-6": {"ms": float(np.median(t_32)) * 1e3, "iterations": info_32.iterations},
        "accuracy": acc,
    }
    return out


def mixed_cpu_check(pa, wl, mg, n=128):
    """CPU leg of the mixed solve: the same generator and settings at n^3, the
    GPU mixed solve against the oracle's pure-FP64 MG-PCG to 1e-12 (all host
    cores). Errors are of the mean-free solutions, relative L-inf."""
    from oracle import oracle as o
    dims = (n, n, n)
    g = pa.Grid3d(dims, np.float64, bc=wl["bc"])
    faces = g.new_simplex_1()
    pa.gen_uniform_faces_device(faces, wl["seed"])
    b = g.new_simplex_2()
    pa.neg_divergence(b, faces)
    x = g.new_simplex_2()
    info = pa.poisson_solve_mixed(x, b, dt=wl["dt"], preconditioner=mg, inner_tol=wl["inner_tol"],
                                  outer_tol=wl["outer_tol"])
    xg, bh = x.download(), b.download()
    P = o.params(dims, bc=wl["bc"], dt=wl["dt"], precond="mg", levels=wl["mg_levels"], tol=1e-12, max_iterations=2000)
    t = time.perf_counter()
    xo, oinfo = o_solve_all_cores(o, P, bh)
    el = time.perf_counter() - t
    ref = xo - xo.mean()
    err = float(np.abs((xg - xg.mean()) - ref).max() / np.abs(ref).max())
    print(f"bench: mixed {dims} vs CPU FP64 oracle ({el:.1f}s, {oinfo.iterations} iterations): error {err:.2e}",
          file=sys.stderr, flush=True)
    return {"dims": list(dims), "err_vs_fp64_cpu": err, "gpu_rel_residual": info.r_linf / info.b_linf,
            "cpu_fp64_iterations": oinfo.iterations, "cpu_fp64_s": el, "cpu_threads": o.omp_lib().or_threads()}


def o_solve_all_cores(o, P, b):
    b = np.array(b, copy=True)
    x = np.zeros_like(b)
    info = o.Info()
    o._fn("or_solve", b.dtype, o.omp_lib())(ctypes.byref(P), o._p(x), o._p(b), ctypes.byref(info))
    return x, info


def batch_step_bytes(dims, iters, levels):
    """Algorithmic bytes of one 2D MG-PCG fluid step (f32 units of N cells:
    advect 6, divergence 3, gradient 5, MG V-cycle 16.25 per level with the
    geometric factor 1/4, CG 2 + 6 + 3 per iteration)."""
    n = dims[0] * dims[1]
    vc = 16.25 * sum(0.25 ** l for l in range(levels - 1))
    return 4.0 * n * (6 + 3 + 5 + vc + (2 + 6 + 3 + vc) * iters)


def bench_batch(args, **_):
    """64 uneven square 2D grids advanced together: one set of launches per
    phase covers every grid, phases replayed as hipGraphs, tile shapes from
    the autotuner. Timed: K steps of all grids from the seeded state."""
    import panopaea_amd as pa
    from panopaea_amd import workloads as W
    wl = W.BATCH64_2D
    grids = W.batch_grids(wl)
    steps = args.steps if args.steps_given else wl["steps"]
    t0 = time.time()
    B = pa.Batch(grids, mg_levels=wl["mg_levels"], mg_nu=wl["mg_nu"], tol=wl["tol"], use_graph=True,
                 autotune=True)
    tiles = B.tiles()
    init = [W.batch_initial(wl, g) for g in grids]
    print(f"bench: batch of {len(grids)} grids set up in {time.time() - t0:.2f}s tiles {tiles}", file=sys.stderr,
          flush=True)

    def load():
        for i, (rho, faces) in enumerate(init):
            B.set_state(i, rho, faces)

    load()
    for k in range(args.warmup):
        its = B.step()
        print(f"bench: warmup {k} max iterations={max(its)}", file=sys.stderr, flush=True)
    load()
    B.ctx.synchronize()
    lat, its_all = [], []
    t_start = time.perf_counter()
    for k in range(steps):
        a = time.perf_counter()
        its_all.append(B.step())
        lat.append(time.perf_counter() - a)
    B.ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    cells = sum(g["dims"][0] * g["dims"][1] for g in grids)
    its = np.array(its_all)
    bytes_total = sum(batch_step_bytes(g["dims"], int(its[s, i]), wl["mg_levels"])
                      for s in range(steps) for i, g in enumerate(grids))
    achieved = bytes_total / elapsed / 1e9
    print(f"bench: {steps} steps {elapsed * 1e3:.1f} ms, CG iterations per step max {its.max(axis=1).tolist()}",
          file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as o
        g = grids[0]
        rho, faces = init[0][0].copy(), init[0][1].copy()
        P = o.params(g["dims"], bc=g["bc"], dt=g["dt"], precond="mg", levels=wl["mg_levels"], tol=wl["tol"],
                     max_iterations=500)
        reps, t = 0, time.perf_counter()
        while time.perf_counter() - t < 5.0:
            o.step(P, rho, faces)
            reps += 1
        el = time.perf_counter() - t
        n0 = g["dims"][0] * g["dims"][1]
        cpu = {"value": n0 * reps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"{reps} steps of grid 0 ({g['dims'][0]}^2) of the batch from its seeded state, CPU "
                         f"oracle (C, f32 fields, MG-PCG), 1 thread, {el:.1f}s"}
    B.close()
    out = {
        "metric": "cell-updates/s, 64 uneven 2D grids advect+project @1e-6 resid, batched hipGraph steps",
        "value": cells * steps / elapsed, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3,
        "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded band-limited velocity, Gaussian density)",
        "config": {"workload": wl["name"], "grids": len(grids), "cells": cells,
                   "sizes": sorted(set(g["dims"][0] for g in grids)),
                   "precond": "multigrid 4 levels, 2+2 red-black GS", "tol": "1e-6 rel L-inf",
                   "tiles": {"natural": list(tiles["nat"]), "red_black": list(tiles["rb"]),
                             "one_pass_2d": ["32x32", "32x16", "16x16", "64x16"][tiles["tb"]]},
                   "cg_iterations_max_per_step": its.max(axis=1).tolist()},
        "roofline": {"bound": "hbm", "kernel": "whole step (model bytes / wall time)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    return out


def serve_run(S, reqs, label):
    """Submits each request at its arrival time, ticks the server, and on
    completion records latency (completion - arrival) and releases the
    request's memory. Returns the latencies, kinds and tick statistics."""
    t0 = S.tick()["now"]
    arrival, lat, kinds = {}, [], []
    cells_done, i, n = 0, 0, len(reqs)
    ticks, active_sum, peak_bytes, peak_active, last_print = 0, 0, 0, 0, 0.0
    while len(lat) < n:
        t = S.tick()
        ticks += 1
        now = t["now"] - t0
        active_sum += t["active"]
        peak_bytes = max(peak_bytes, t["bytes_used"])
        peak_active = max(peak_active, t["active"])
        for rid in S.completed():
            st = S.status(rid)
            ta, nd = arrival.pop(rid)
            lat.append(st["t_done"] - t0 - ta)
            kinds.append(nd)
            cells_done += st["cells"]
            if not st["converged"]:
                print(f"bench: request {rid} did not converge", file=sys.stderr, flush=True)
            S.release(rid)
        while i < n and reqs[i][0] <= now:
            ta, g, seed, nm, km = reqs[i]
            rid = S.submit(g, seed=seed, n_modes=nm, k_max=km)
            arrival[rid] = (ta, len(g["dims"]))
            i += 1
        if now - last_print >= 10.0:
            last_print = now
            print(f"bench: {label} t={now:.0f}s submitted {i} done {len(lat)} active {t['active']} queued "
                  f"{t['queued']} bytes {t['bytes_used'] / 2**30:.2f} GiB", file=sys.stderr, flush=True)
        if t["active"] == 0 and t["queued"] == 0 and i < n and reqs[i][0] > now:
            time.sleep(min(reqs[i][0] - now, 0.01))
    elapsed = S.tick()["now"] - t0
    return dict(lat_ms=np.array(lat) * 1e3, kinds=np.array(kinds), cells=cells_done, elapsed=elapsed, ticks=ticks,
                mean_active=active_sum / max(ticks, 1), peak_active=peak_active, peak_bytes=peak_bytes)


def bench_serve(args, **_):
    """Open-loop serving: projection requests arrive by a seeded Poisson
    process (200/s for 60 s); velocity fields are generated on the device at
    admission from each request's seed. Then the first 600 requests of the
    same stream twice more: arriving ten times faster (overlapping requests:
    continuous batching under load), and one at a time (each request alone),
    so the overlapped throughput stands beside the one-at-a-time rate."""
    import panopaea_amd as pa
    from panopaea_amd import workloads as W
    wl = W.SERVE
    reqs = W.serve_requests(wl, args.duration)
    S = pa.Server(hbm_budget=64 << 30, max_active=1024, mg_levels=wl["mg_levels"], mg_nu=wl["mg_nu"], tol=wl["tol"],
                  iters_per_tick=2)
    # warm-up: one request of each extreme size, not measured
    for g in (dict(dims=(1024, 1024), dt=0.5 / 1024), dict(dims=(256, 256, 256), dt=0.5 / 256, bc=(0, 0, 0, 0, 0, 1))):
        S.submit(g, seed=12345)
    S.run_until_idle(lambda rid: S.release(rid))
    print(f"bench: serve {len(reqs)} requests over {reqs[-1][0]:.1f}s", file=sys.stderr, flush=True)
    r = serve_run(S, reqs, "serve")
    lat_ms, kinds = r["lat_ms"], r["kinds"]
    # overlap: the first 600 requests at ten times the rate, then each of them alone
    sub = reqs[:600]
    fast = [(q[0] / 10.0,) + tuple(q[1:]) for q in sub]
    b = serve_run(S, fast, "burst")
    t = time.perf_counter()
    cells_alone = 0
    for _, g, seed, nm, km in sub:
        S.submit(g, seed=seed, n_modes=nm, k_max=km)
        S.run_until_idle(lambda rid: S.release(rid))
        cells_alone += int(np.prod(g["dims"]))
    alone_s = time.perf_counter() - t
    S.close()
    sub_cells = sum(int(np.prod(g["dims"])) for _, g, *_ in sub)
    print(f"bench: burst of {len(sub)} requests: {b['elapsed']:.2f}s, mean active {b['mean_active']:.1f}, "
          f"alone {alone_s:.2f}s", file=sys.stderr, flush=True)
    total_cells = sum(int(np.prod(g["dims"])) for _, g, *_ in reqs)
    out = {
        "metric": "served projection solves: p50/p99 latency (ms) and cells/s, Poisson arrivals 200/s, mixed 2D/3D",
        "value": r["cells"] / r["elapsed"], "unit": "cells/s", "n_gpus": 1, "steps": len(reqs), "warmup": 2,
        "ms_per_step": float(np.mean(lat_ms)), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "p50_ms_2d": float(np.percentile(lat_ms[kinds == 2], 50)), "p99_ms_2d": float(np.percentile(lat_ms[kinds == 2], 99)),
        "p50_ms_3d": float(np.percentile(lat_ms[kinds == 3], 50)), "p99_ms_3d": float(np.percentile(lat_ms[kinds == 3], 99)),
        "dtype": "f32", "data": "synthetic (seeded arrivals, sizes and band-limited velocity)",
        "config": {"workload": wl["name"], "requests": len(reqs), "arrival_rate": wl["rate"],
                   "duration_s": reqs[-1][0], "offered_cells_per_s": total_cells / reqs[-1][0],
                   "elapsed_s": r["elapsed"], "ticks": r["ticks"], "mean_active": r["mean_active"],
                   "peak_active": r["peak_active"], "peak_hbm_gib": r["peak_bytes"] / 2**30, "hbm_budget_gib": 64,
                   "precond": "multigrid 4 levels, 2+2 red-black GS", "tol": "1e-6 rel L-inf"},
        "overlap": {"requests": len(sub), "arrival_rate": wl["rate"] * 10, "elapsed_s": b["elapsed"],
                    "cells_per_s": b["cells"] / b["elapsed"], "mean_active": b["mean_active"],
                    "peak_active": b["peak_active"], "p50_ms": float(np.percentile(b["lat_ms"], 50)),
                    "p99_ms": float(np.percentile(b["lat_ms"], 99)),
                    "one_at_a_time_s": alone_s, "one_at_a_time_cells_per_s": sub_cells / alone_s,
                    "note": "the same 600 requests arriving 10x faster (their work overlaps on the device) "
                            "against each served alone"},
        "roofline": None, "cpu_baseline": None,
    }
    return out


def bench_stream_copy(args):
    """Device copy of a 1 GiB f32 buffer (torch's elementwise copy kernel):
    the measured copy bandwidth, and the known-size kernel that calibrates
    the FETCH_SIZE / WRITE_SIZE counters (1 GiB read, 1 GiB written per
    launch)."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(1, args.warmup)):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2.0 * 4 * n / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "bytes_read": 4 * n,
                      "bytes_written": 4 * n, "ms_per_copy": ms, "steps": args.steps}), flush=True)


def profiled_traffic(kernel, dims):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE counter summary of the same workload (None if there is none)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_smoke3d_{'x'.join(map(str, dims))}_pmc.md")
    if tuple(dims) == (256, 256, 256):
        path = os.path.join(ROOT, "profiles", f"{kernel}_smoke3d_256_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, dims, cfg, init_vel):
    """The CPU leg: the oracle (a C restatement of the reference method) on
    one step of the same workload from the same initial state, on one thread
    and on all host cores (the OpenMP build of the same sources, standing in
    for the reference's rayon loops); and, as the checker, the GPU step's
    error against the FP64 oracle on a bounded sample of the workload."""
    from oracle import oracle as o
    P = o.params(dims, bc=wl["bc"], dt=cfg.dt, precond="mg", levels=wl["mg_levels"], tol=wl["tol"],
                 max_iterations=500, buoyancy=wl["buoyancy"], src_center=cfg.src_center, src_radius=cfg.src_radius)
    n = int(np.prod(dims))
    legs = {}
    for name, L in (("one_thread", None), ("all_cores", o.omp_lib())):
        rho = np.zeros(n, np.float32)
        faces = init_vel.copy()
        t = time.perf_counter()
        _, info = o.step(P, rho, faces, L)
        el = time.perf_counter() - t
        threads = (L or o.lib).or_threads()
        print(f"bench: cpu oracle step ({name}, {threads} threads) {el:.2f}s iterations={info.iterations}",
              file=sys.stderr, flush=True)
        legs[name] = {"value": n / el, "unit": "cell-updates/s", "cores": threads, "seconds": el,
                      "cg_iterations": info.iterations}
    best = legs["all_cores"]
    cpu = {"value": best["value"], "unit": "cell-updates/s", "cores": best["cores"], "kind": "port",
           "sample": f"step 1 of the {'x'.join(map(str, dims))} smoke workload from the same seeded initial state, "
                     f"CPU oracle (C, gcc -O3, f32 fields, MG-PCG) on all {best['cores']} threads, "
                     f"{best['seconds']:.1f}s; one thread: {legs['one_thread']['value']:.3g} cell-updates/s "
                     f"({legs['one_thread']['seconds']:.1f}s)",
           "legs": legs, "nproc": os.cpu_count(), "cpu_model": cpu_model()}
    return cpu, fp64_check(wl, o)


def fp64_check(wl, o, n=64, steps=5):
    """The f32 GPU step against the FP64 oracle on a bounded sample: the same
    workload at n^3 (same generator, source, buoyancy, MG-PCG to 1e-6),
    `steps` steps from the same seeded state; relative L-inf errors of the
    velocity and the density per step, and the projected velocity's
    divergence against ||b||_inf for both."""
    import panopaea_amd as pa
    from panopaea_amd import diagnostics as D
    from panopaea_amd import workloads as W
    w = dict(wl, dims=(n, n, n))
    g = W.make_grid(w)
    cfg = W.step_config(w)
    rho, faces = W.initial_state(w, g)
    P = o.params(w["dims"], bc=w["bc"], dt=cfg.dt, precond="mg", levels=w["mg_levels"], tol=w["tol"],
                 max_iterations=500, buoyancy=w["buoyancy"], src_center=cfg.src_center, src_radius=cfg.src_radius)
    r0, v0 = g.new_simplex_2().upload(rho), g.new_simplex_1().upload(faces)
    r1, v1, pr = g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_2()
    r64, f64 = rho.astype(np.float64), faces.astype(np.float64)
    ev, er, first = [], [], None
    for k in range(steps):
        info = pa.fluid_step(g, cfg, r0, v0, r1, v1, pr)
        p64, i64 = o.step(P, r64, f64, o.omp_lib())
        gv, gr = v1.download(), r1.download()
        ev.append(float(np.abs(gv - f64).max() / np.abs(f64).max()))
        er.append(float(np.abs(gr - r64).max() / np.abs(r64).max()))
        if k == 0:
            d32 = D.projection_report(g, gv, pr.download(), cfg.dt, info.b_linf)
            d64 = D.projection_report(g, f64, p64, cfg.dt, i64.b_linf)
            first = {"gpu_f32_div_rel_linf": d32["div_rel_linf"], "fp64_div_rel_linf": d64["div_rel_linf"]}
        r0, r1, v0, v1 = r1, r0, v1, v0
    out = {"sample": f"{n}^3 smoke workload, {steps} steps from the seeded state", "velocity_rel_linf": ev,
           "density_rel_linf": er, "step0": first}
    print(f"bench: fp64 check {out}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
