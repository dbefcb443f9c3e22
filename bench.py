#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of the fused Euler step on one MI355X.

Workload (BASELINE config "C4"): 4096 x 2048 cells on [0,2] x [0,1], Mach 3
inflow (rho, u, v, p) = (1.4, 3, 0, 1) past 9 seeded bodies (seed 20261202:
circles and convex polygons, exact signed-distance level set), MUSCL + van Leer
limiter + HLL flux, IDW ghost-fluid boundary, supersonic inflow / zero-gradient
edges, FP64. A step = one fused step kernel (reconstruction, fluxes, update,
edge ring, dt and mass reductions) + ghost fill + edge kernel. The state is
initialised on the device before the timed region; dt never leaves the device.

Prints one JSON line. `value` = N_x*N_y*steps / t_loop (HIP events around K
queued steps on the solve's stream). `roofline` is for the step kernel alone:
65 algorithmic bytes per cell update (read U^n 4xFP64, write U^{n+1} 4xFP64,
one uint8 cell flag) divided by its mean launch duration (HIP events around each
launch in a second timed pass). `cpu_baseline` times the C++ oracle (the
reference's own loop structure, OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL = 65.0
HBM_PEAK_GBS = 8000.0


def gpu_leg(args):
    import ibtfo_amd as ib
    from ibtfo_amd import workloads as W
    cfg = W.c4_config(ib, N_x=args.nx, N_y=args.ny, Riemann=args.riemann, Limiter=args.limiter,
                      Boundary=args.boundary)
    s = ib.Solver(cfg)
    s.geometry("shapes", W.c4_shapes())
    s.initialize("uniform", W.C4_INFLOW)
    s.prepare(True)
    nghost = s.num_ghosts()
    if args.warmup:
        s.step(args.warmup)
        s.status()
    total_ms, kern_ms = s.time_steps(args.steps, kernel=True)
    it, t, dt, rc = s.status()
    cells = args.nx * args.ny
    ms_step = total_ms / args.steps
    value = cells * args.steps / (total_ms / 1e3)
    achieved = BYTES_PER_CELL * cells / (kern_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c4.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = dict(value=value, ms_per_step=ms_step, kernel_ms=kern_ms, achieved=achieved, traffic=traffic,
               iter=it, t=t, dt=dt, status=rc, nghost=nghost)
    s.close()
    return out


def cpu_leg(args):
    """The oracle (C++ restatement of the reference loop, OpenMP) on a bounded
    sample: the same workload on a 512 x 256 grid for a fixed number of steps."""
    import ctypes as C
    from oracle import oracle as O
    from ibtfo_amd import workloads as W
    nx, ny = 512, 256
    steps = args.cpu_steps
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    O.lib().or_set_num_threads(threads)
    case = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=10 ** 9,
                  CFL_number=0.5, Scheme="MUSCL", Riemann=args.riemann, Limiter=args.limiter,
                  Boundary=args.boundary, case="uniform", edge="inflow", quirks=0,
                  extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    M, N = case.M, case.N
    dx = 2.0 / nx
    Li = -dx * 2 + 0.0 + dx / 2
    x = (Li + np.arange(M) * dx)[None, :].repeat(N, 0)
    y = (Li + np.arange(N) * dx)[:, None].repeat(M, 1)
    case.phi = W.shapes_phi(W.c4_shapes(), x, y).reshape(-1)
    st = O.setup(case)
    t0 = time.perf_counter()
    r = O.run_oracle(case, steps, state=st)
    el = time.perf_counter() - t0
    return dict(value=nx * ny * r["steps"] / el, cores=threads, steps=r["steps"], seconds=el,
                sample="%dx%d cells, %d steps of the same workload (oracle, OpenMP %d threads)" % (
                    nx, ny, r["steps"], threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--riemann", default="HLL")
    ap.add_argument("--limiter", default="VanLeer")
    ap.add_argument("--boundary", default="IDWBaseGFM")
    ap.add_argument("--cpu-steps", type=int, default=600)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the whole domain on one GPU; --gpus %d is not supported "
              "(multi-GPU slab decomposition is a design note only)" % args.gpus, file=sys.stderr)
        return 2
    g = gpu_leg(args)
    cpu = None
    if not args.no_cpu:
        c = cpu_leg(args)
        cpu = dict(value=c["value"], unit="cell-updates/s", cores=c["cores"], kind="port", sample=c["sample"])
    line = {
        "metric": "cell-updates/s (frac. HBM roofline); batched p50/p99 solve latency at fixed L1",
        "value": g["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Mach 3 inflow past 9 seeded bodies (seed 20261202), %dx%d cells on [0,2]x[0,1], "
                               "MUSCL + %s + %s, %s, inflow/zero-gradient edges, FP64"
                               % (args.nx, args.ny, args.limiter, args.riemann, args.boundary),
                   "N_x": args.nx, "N_y": args.ny, "ghost_cells": g["nghost"], "steps_taken": g["iter"],
                   "parallelism": "single GPU, whole domain resident"},
        "roofline": {"bound": "hbm", "achieved": g["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": g["achieved"] / HBM_PEAK_GBS, "traffic": g["traffic"],
                     "kernel": "k_step<MUSCL,HLL,VanLeer>", "kernel_ms": g["kernel_ms"],
                     "bytes_per_cell": BYTES_PER_CELL},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
