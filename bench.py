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
during the timed steps), and the CPU oracle timed on a bounded sample of the
same workload on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel smooth_l0] [--no-cpu]
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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="smooth_l0", help="kernel whose roofline is reported")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    ap.add_argument("--dims", default="256,256,256")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the fluid step runs on a single MI355X); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)

    import panopaea_amd as pa
    from panopaea_amd import _lib
    from panopaea_amd import workloads as W

    wl = W.SMOKE3D_256
    dims = tuple(int(v) for v in args.dims.split(","))
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
        roof = {"bound": "hbm", "kernel": args.kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "launches": launches.value, "avg_us": round(avg_s * 1e6, 2),
                "bytes_per_launch": bpl.value}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, dims, cfg, init_vel)

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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(wl, dims, cfg, init_vel):
    """The CPU oracle (a C restatement of the reference method, single
    thread) on one step of the same workload from the same initial state."""
    from oracle import oracle as o
    P = o.params(dims, bc=wl["bc"], dt=cfg.dt, precond="mg", levels=wl["mg_levels"], tol=wl["tol"],
                 max_iterations=500, buoyancy=wl["buoyancy"], src_center=cfg.src_center, src_radius=cfg.src_radius)
    rho = np.zeros(int(np.prod(dims)), np.float32)
    faces = init_vel.copy()
    t = time.perf_counter()
    _, info = o.step(P, rho, faces)
    el = time.perf_counter() - t
    print(f"bench: cpu oracle step {el:.2f}s iterations={info.iterations}", file=sys.stderr, flush=True)
    return {"value": int(np.prod(dims)) / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"step 1 of the {'x'.join(map(str, dims))} smoke workload from the same seeded initial "
                      f"state, CPU oracle (C, -O2, f32 fields, MG-PCG), 1 thread, {el:.1f}s"}


if __name__ == "__main__":
    main()
