#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of the Euler step on one MI355X.

Workload: 4096 x 2048 cells on [0,2] x [0,1], Mach 3 inflow (rho, u, v, p) =
(1.4, 3, 0, 1) past 9 seeded bodies (seed 20261202: circles and convex
polygons) given as a binary mask image (one pixel per cell) in a PNG file, which
the library decodes and whose level set it builds on the device (mask -> edge
pixels -> signed distance), cut-cell boundary
with small-cell merging, MUSCL + van Leer limiter, HLL flux (the headline) and
Steger-Warming flux (second leg), supersonic inflow / zero-gradient edges, FP64;
a third leg runs the reference's IDW ghost-fluid boundary in place of the cut
cells.
A step = the step kernel (reconstruction, fluxes, update, per-block dt/mass
partials), the small-cell merge kernel (with the merged cells' dt/mass
partials), the dt/clock kernel, and the ghost-fill kernels (the first of them
also fills the edge ring). The state is initialised on the device before the timed region and
dt never leaves the device. The step kernel's tile height is picked by the
library's autotuner; the timed loop replays a captured HIP graph of the step
sequence.

Prints one JSON line. `value` = N_x*N_y*steps / t_loop (HIP events around K
steps on the solve's stream). `roofline` is for the step kernel alone: 65
algorithmic bytes per cell update (read U^n 4xFP64, write U^{n+1} 4xFP64, one
uint8 cell flag) divided by its mean launch duration (HIP events around each
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
# the CPU legs' OpenMP threads stay on neighbouring cores (read when the OpenMP
# runtime starts, with the first load of the oracle library)
os.environ.setdefault("OMP_PROC_BIND", "close")
# The process's CPU share, read before any OpenMP runtime starts: with
# OMP_PROC_BIND set, the runtime binds the initial thread to the first place, and
# the main thread's affinity mask then names that one CPU.
AFFINITY = sorted(os.sched_getaffinity(0))

def _say(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


BYTES_PER_CELL = 65.0
HBM_PEAK_GBS = 8000.0
STRIP_COLUMNS = 60  # interior columns per step-kernel strip (ibtfo_tile's grid: ceil(N_x / 60) strips across)


def traffic_file(riemann):
    """profiles/ summary of the HBM counter passes for a flux's step kernel."""
    return "traffic_step_mach3_mask_cutcell%s.json" % ("" if riemann in (None, "HLL") else "_" + riemann.lower())


def workload_name(args):
    geo = ("binary mask read from a %dx%d PNG, level set built on the device" % (args.nx + 4, args.ny + 4)
           if args.geometry == "mask"
           else "exact signed-distance level set")
    return ("Mach 3 inflow past 9 seeded bodies (seed 20261202; %s), %dx%d cells on [0,2]x[0,1], "
            "MUSCL + %s + %s, %s, inflow/zero-gradient edges, FP64"
            % (geo, args.nx, args.ny, args.limiter, args.riemann, args.boundary))


_MASK = {}


def workload_geometry(args, s):
    """The bodies on a solver: the binary mask, rasterised once into a PNG file
    that the library reads (its PNG decoder, then the level set on the device),
    or the exact signed distance of the same shapes."""
    from ibtfo_amd import workloads as W
    if args.geometry == "mask":
        key = (args.nx, args.ny)
        if key not in _MASK:
            import tempfile
            path = os.path.join(tempfile.mkdtemp(prefix="ibtfo_mask_"), "bodies_%dx%d.png" % key)
            W.write_png(path, W.bodies_mask(args.nx, args.ny))
            _MASK[key] = path
        s.geometry_png(_MASK[key])
    else:
        s.geometry("shapes", W.bodies_shapes())


def gpu_leg(args, precision="FP64", keep_state=False, riemann=None, boundary=None, keep_phi=False):
    import ibtfo_amd as ib
    from ibtfo_amd import workloads as W
    cfg = W.mach3_bodies_config(ib, N_x=args.nx, N_y=args.ny, Riemann=riemann or args.riemann,
                                Limiter=args.limiter, Boundary=boundary or args.boundary)
    cfg.precision = ib.PRECISION[precision]
    s = ib.Solver(cfg)
    if args.geometry == "mask":
        workload_geometry(args, s)  # rasterises the mask into a PNG once (host); untimed
        s.close()
        s = ib.Solver(cfg)
    t_geo = time.perf_counter()
    workload_geometry(args, s)  # PNG decode, mask upload, level set, normals, cut cells, ghost list
    geometry_ms = (time.perf_counter() - t_geo) * 1e3
    png_decode_ms = None
    if args.geometry == "mask":  # the host part of it: one decode of the PNG into a ready buffer
        import ctypes as C
        img = np.empty((args.ny + 4, args.nx + 4, 4), np.uint8)
        w, h = C.c_int(), C.c_int()
        t_png = time.perf_counter()
        rc = ib.lib().ibtfo_png_decode(os.fsencode(_MASK[(args.nx, args.ny)]), img.ctypes.data_as(C.c_void_p),
                                       C.c_longlong(img.size), C.byref(w), C.byref(h))
        png_decode_ms = (time.perf_counter() - t_png) * 1e3 if rc == ib.OK else None
    s.initialize("uniform", W.MACH3_INFLOW)
    s.prepare(True)
    nghost = s.num_ghosts()
    ncut, ngroups = s.cut_counts()
    tuned = None
    if args.rows > 0:
        s.set_rows(args.rows)
    elif not args.no_tune:
        tuned = s.autotune(args.tune_steps)
    rows, nbx, nby = s.tile()
    form = "row-parallel" if s.tile_form() else "strip"
    if args.warmup:
        s.step(args.warmup)
        s.status()
    # K timed steps in 5 equal slices (HIP events around each on the solve's
    # stream): value = all K steps, plus the slices' median
    cells = args.nx * args.ny
    nsl = 5 if args.steps >= 5 else 1
    slices = [args.steps // nsl] * nsl
    slices[-1] += args.steps - sum(slices)
    chunk = 0
    if args.graph_chunk > 0:
        # capture the step sequence once, `chunk` steps long, a divisor of the
        # slice length: every timed step is replayed from the graph
        chunk = min(args.graph_chunk, slices[0])
        while slices[0] % chunk or slices[-1] % chunk:
            chunk -= 1
        s.step(chunk, graph_chunk=chunk)
        s.status()
    slice_ms = [s.time_steps(n, kernel=False)[0] for n in slices]
    total_ms = sum(slice_ms)
    slice_rates = sorted(cells * n / (m / 1e3) for n, m in zip(slices, slice_ms))
    _, kern_ms = s.time_steps(min(args.steps, args.kernel_steps), kernel=True)
    it, t, dt, rc = s.status()
    ms_step = total_ms / args.steps
    value = cells * args.steps / (total_ms / 1e3)
    achieved = BYTES_PER_CELL * cells / (kern_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", traffic_file(riemann or args.riemann))
    if os.path.exists(tpath) and precision == "FP64" and boundary is None:
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("workload") == workload_name(argparse.Namespace(**dict(vars(args), riemann=riemann or args.riemann))):
            traffic = tr.get("bytes_per_launch")
    out = dict(value=value, ms_per_step=ms_step, kernel_ms=kern_ms, achieved=achieved, traffic=traffic,
               iter=it, t=t, dt=dt, status=rc, nghost=nghost, ncut=ncut, ngroups=ngroups, rows=rows,
               slice_rates=slice_rates, median=slice_rates[len(slice_rates) // 2], geometry_ms=geometry_ms, png_decode_ms=png_decode_ms, grid=[nbx, nby],
               form=form,
               tuned=tuned is not None, chunk=chunk, slices=slices)
    if keep_state:
        out["rho"] = s.state()[0::4].copy()
        out["types"] = s.xycoord()[5::6].copy()
    if keep_phi:
        out["phi"] = s.xycoord()[2::6].copy()
    s.close()
    return out


def _physical_cores():
    """Physical cores of the machine as lscpu reports them (cores per socket x
    sockets), or None."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        return None
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        return int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (KeyError, ValueError):
        return None


def cpu_leg(args, phi):
    """The oracle (C++ restatement of the reference loop, OpenMP) on the same
    workload at its full size (4096 x 2048, the device's mask level set handed
    over as a per-cell field), truncated to its first `cpu_steps` steps; and
    on one thread for its first `cpu_steps_1t` steps. Threads: this process's
    CPU share (its affinity mask, and OMP_NUM_THREADS where the environment
    sets it) within the machine's physical cores; ACC_NUM_CORES of the
    reference's multicore build would be the same number."""
    from oracle import oracle as O
    nx, ny = args.nx, args.ny
    affinity = AFFINITY
    phys = _physical_cores()
    threads = _cpu_threads(args)
    O.lib().or_set_num_threads(threads)
    case = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=10 ** 9,
                  CFL_number=0.5, Scheme="MUSCL", Riemann=args.riemann, Limiter=args.limiter,
                  Boundary=args.boundary, case="uniform", edge="inflow", quirks=0,
                  extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    case.phi = phi
    st = O.setup(case)
    t0 = time.perf_counter()
    r = O.run_oracle(case, args.cpu_steps, state=st)
    el = time.perf_counter() - t0
    _say("CPU baseline: %d steps on %d threads in %.1f s" % (r["steps"], threads, el))
    del st, r["U_OLD"], r["U_NEW"]
    O.lib().or_set_num_threads(1)
    st1 = O.setup(case)
    t1 = time.perf_counter()
    r1 = O.run_oracle(case, args.cpu_steps_1t, state=st1)
    el1 = time.perf_counter() - t1
    del st1
    O.lib().or_set_num_threads(threads)
    cpus = "%d-%d" % (affinity[0], affinity[-1]) if affinity == list(range(affinity[0], affinity[-1] + 1)) else \
        ",".join(map(str, affinity))
    return dict(value=nx * ny * r["steps"] / el, cores=threads, steps=r["steps"], seconds=el,
                physical_cores=phys, affinity=cpus, acc_num_cores=threads,
                sample="the workload at full size (%dx%d cells), its first %d steps (oracle, OpenMP %d threads, "
                       "OMP_PROC_BIND=%s, CPUs %s of a machine with %s physical cores)"
                % (nx, ny, r["steps"], threads, os.environ.get("OMP_PROC_BIND"), cpus, phys),
                value_1t=nx * ny * r1["steps"] / el1, steps_1t=r1["steps"], seconds_1t=el1)


def copy_bandwidth():
    """Device-to-device copy rate on this GPU (GB/s, bytes read + written): a
    1 GiB hipMemcpyAsync, 10 repetitions between HIP events, through the HIP
    runtime the library itself is linked against (libamdhip64.so.7)."""
    import ctypes as C
    import ibtfo_amd as ib
    ib.lib()  # loads the runtime
    hip = C.CDLL("libamdhip64.so.7")
    n = 1 << 30
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    gbs = None
    if ok and hip.hipMemset(a, 1, C.c_size_t(n)) == 0 and hip.hipEventCreate(C.byref(e0)) == 0 and \
            hip.hipEventCreate(C.byref(e1)) == 0:
        D2D = 3  # hipMemcpyDeviceToDevice
        hip.hipMemcpyAsync(b, a, C.c_size_t(n), D2D, None)
        hip.hipEventRecord(e0, None)
        for _ in range(10):
            hip.hipMemcpyAsync(b, a, C.c_size_t(n), D2D, None)
        hip.hipEventRecord(e1, None)
        ms = C.c_float()
        if hip.hipEventSynchronize(e1) == 0 and hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0:
            gbs = 10 * 2 * n / (ms.value / 1e3) / 1e9
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    return gbs


def _cfg_frac(cells, steps, ms, stages=1):
    """Fraction of the HBM roofline of a whole run: 65 algorithmic bytes per cell
    and step-kernel stage (RK2: two), over the run's time, against 8 TB/s."""
    return BYTES_PER_CELL * cells * stages * steps / (ms / 1e3) / 1e9 / HBM_PEAK_GBS


def _cpu_threads(args):
    """Threads for the CPU legs: this process's CPU share (affinity mask, and
    OMP_NUM_THREADS where the environment sets it), within the physical cores."""
    affinity = len(AFFINITY)
    t = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or affinity
    phys = _physical_cores()
    return min(t, affinity, phys or t)


def config_legs(args, legs=("sod", "channel", "vortex")):
    """The reference's other configurations, the GPU beside the CPU oracle: the
    400x100 Sod tube (MHM + minmod + HLLC, to t = 0.2), the inclined Sod channel
    (800x400, MUSCL + minmod + HLL, IDW ghost fluid, to t = 0.2) and the
    isentropic vortex (1024^2, periodic, MUSCL + van Leer + HLL with Heun's RK2,
    to t = 10, FP64 and mixed, L1 density error against the exact solution).
    GPU: HIP events around the whole run (graph-replayed steps); CPU: the oracle
    (OpenMP) on the whole run, or its first `cpu_steps` steps where stated, and,
    for the two configurations the reference itself can run, its own
    translation units built here (`cpu_reference`)."""
    import ibtfo_amd as ib
    from ibtfo_amd import workloads as W
    from oracle import oracle as O
    threads = _cpu_threads(args)
    O.lib().or_set_num_threads(threads)
    out = []

    def gpu_run(cfg, ls, init, apply_boundary, chunk=None):
        chunk = args.graph_chunk if chunk is None else chunk  # 0: plain launches
        s = ib.Solver(cfg)
        s.geometry(*ls)
        s.initialize(*init)
        s.prepare(apply_boundary)
        n = s.run(graph_chunk=chunk)  # untimed: how many steps the run takes
        s.close()
        s = ib.Solver(cfg)
        s.geometry(*ls)
        s.initialize(*init)
        s.prepare(apply_boundary)
        rows, _ = s.autotune(min(20, max(1, n // 8)))  # probes restore the state and clock
        s.step(0, graph_chunk=chunk)  # capture only
        ms, _ = s.time_steps(n, kernel=False)
        it, t, dt, rc = s.status()
        s.tuned_rows = rows
        s.tuned_form = "row-parallel" if s.tile_form() else "strip"
        s.tuned_one_kernel = s.one_kernel_steps()  # the deferred finish: one kernel per step
        return s, n, ms, t, rc

    def cpu_run(case, nsteps):
        st = O.setup(case)
        t0 = time.perf_counter()
        r = O.run_oracle(case, nsteps, state=st)
        return r["steps"], time.perf_counter() - t0

    def ref_run(case, nsteps, cells, note):
        """The reference's own translation units (oracle/_ref, built from its
        sources with g++ -fopenacc: OpenACC's host fallback, one thread), with
        their console output sent to stderr; None where they are not built."""
        if not O.ref_available(case):
            return None
        import ctypes as C
        libc = C.CDLL(None)
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            t0 = time.perf_counter()
            r = O.run_ref(case, nsteps)
            el = time.perf_counter() - t0
        finally:
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        return dict(value=cells * r["steps"] / el, steps=r["steps"], cores=1, kind="reference", note=note)

    def sod_leg():
        # 400x100 Sod tube
        cfg = ib.config(N_x=400, N_y=100, hi_x=1.0, hi_y=0.25, Psy_time=0.2, max_iter=10 ** 9, CFL_number=0.5,
                        Scheme="MHM", Riemann="HLLC", Limiter="minmod", Boundary="Regular")
        s, n, ms, t, rc = gpu_run(cfg, ("none", ()), ("sod_x", (0.5, 1, 0, 0, 1, 0.125, 0, 0, 0.1)), False)
        s.close()
        c = O.Case(N_x=400, N_y=100, hi_y=0.25, Scheme="MHM", Riemann="HLLC", Limiter="minmod", case="sodx",
                   Psy_time=0.2)
        cn, cs = cpu_run(c, 10 ** 6)
        c.Riemann = "HLL"  # the reference's HLLC directory holds HLL's code; its build is the HLL one
        ref = ref_run(c, 10 ** 6, 400 * 100, "the whole run; the x-split Sod hook compiled into the reference's loop")
        _say("Sod tube measured")
        return dict(workload="Sod tube 400x100 on [0,1]x[0,0.25], MHM + minmod + HLLC, Regular, to t = 0.2",
                    steps=n, gpu_value=400 * 100 * n / (ms / 1e3), gpu_ms=ms, gpu_us_per_step=1e3 * ms / n,
                    roofline_frac=_cfg_frac(400 * 100, n, ms), tile_rows=s.tuned_rows, tile_form=s.tuned_form, one_kernel_steps=s.tuned_one_kernel,
                    status=rc,
                    cpu_value=400 * 100 * cn / cs, cpu_steps=cn, cpu_cores=threads, cpu_reference=ref)

    def channel_leg():
        # inclined Sod channel, 800x400, IDW
        cfg = ib.config(N_x=800, N_y=400, hi_x=2.0, hi_y=1.0, Psy_time=0.2, max_iter=10 ** 9, CFL_number=0.5,
                        Scheme="MUSCL", Riemann="HLL", Limiter="minmod", Boundary="IDWBaseGFM")
        s, n, ms, t, rc = gpu_run(cfg, ("channel", (30.0, 0.4)),
                                  ("incline", (30.0, 1, 0, 0, 1, 0.125, 0, 0, 0.1)), True)
        s.close()
        c = O.Case(N_x=800, N_y=400, hi_x=2.0, hi_y=1.0, Boundary="IDWBaseGFM", case="incline", Psy_time=0.2)
        cn, cs = cpu_run(c, args.cpu_incline_steps)
        ref = ref_run(c, args.ref_incline_steps, 800 * 400, "the first %d steps of the run" % args.ref_incline_steps)
        _say("inclined channel measured")
        return dict(workload="inclined Sod channel 800x400 on [0,2]x[0,1] (30 deg, AxisD 0.4), MUSCL + minmod "
                             "+ HLL, IDW ghost fluid, to t = 0.2",
                    steps=n, gpu_value=800 * 400 * n / (ms / 1e3), gpu_ms=ms, gpu_us_per_step=1e3 * ms / n,
                    roofline_frac=_cfg_frac(800 * 400, n, ms), tile_rows=s.tuned_rows, tile_form=s.tuned_form, one_kernel_steps=s.tuned_one_kernel,
                    status=rc,
                    cpu_value=800 * 400 * cn / cs, cpu_steps=cn, cpu_cores=threads,
                    cpu_note="first %d steps of the run" % cn, cpu_reference=ref)

    def vortex_leg():
        # isentropic vortex 1024^2, FP64 and mixed
        nv = 1024
        h = 10.0 / nv
        xc = (np.arange(nv) + 0.5) * h
        X, Y = np.meshgrid(xc, xc)
        vort = {}
        for prec in ["FP64", "mixed"]:
            cfg = ib.config(N_x=nv, N_y=nv, hi_x=10.0, hi_y=10.0, Psy_time=10.0, max_iter=10 ** 9, CFL_number=0.5,
                            Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="Regular", Time="RK2",
                            Edge="periodic", quirks=0, precision=ib.PRECISION[prec])
            s, n, ms, t, rc = gpu_run(cfg, ("none", ()), ("shu", (5.0, 5.0, 5.0, 1.0, 1.0, 1.0, 1.0)), False)
            rho = s.state().reshape(nv + 4, nv + 4, 4)[2:-2, 2:-2, 0]
            s.close()
            vort[prec] = dict(steps=n, ms=ms, t=t, rc=rc, rows=s.tuned_rows, form=s.tuned_form,
                              l1=float(np.mean(np.abs(rho - W.shu_exact(X, Y, t)))), rho=rho)
        _say("isentropic vortex on the GPU measured")
        c = O.Case(N_x=nv, N_y=nv, hi_x=10.0, hi_y=10.0, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer",
                   Time="RK2", case="shu", edge="periodic", quirks=0, Psy_time=10.0)
        cn, cs = cpu_run(c, args.cpu_vortex_steps)
        f, m = vort["FP64"], vort["mixed"]
        return dict(workload="isentropic vortex 1024x1024 on [0,10]^2 (strength 5, free stream (1,1,1,1)), "
                             "periodic, MUSCL + VanLeer + HLL, RK2 (Heun), to t = 10",
                    steps=f["steps"], gpu_value=nv * nv * f["steps"] / (f["ms"] / 1e3), gpu_ms=f["ms"],
                    gpu_us_per_step=1e3 * f["ms"] / f["steps"],
                    roofline_frac=_cfg_frac(nv * nv, f["steps"], f["ms"], stages=2), tile_rows=f["rows"], tile_form=f["form"],
                    status=f["rc"], mixed_value=nv * nv * m["steps"] / (m["ms"] / 1e3),
                    l1_rho_exact_fp64=f["l1"], l1_rho_exact_mixed=m["l1"], l1_ratio=m["l1"] / f["l1"],
                    l1_rho_mixed_vs_fp64=float(np.mean(np.abs(m["rho"] - f["rho"]))),
                    cpu_value=nv * nv * cn / cs, cpu_steps=cn, cpu_cores=threads,
                    cpu_note="first %d steps of the run" % cn)

    for name, leg in (("sod", sod_leg), ("channel", channel_leg), ("vortex", vortex_leg)):
        if name in legs:
            out.append(leg())
    return out


def _plain_solve(it):
    from ibtfo_amd import serving as S
    return S.plain_solve(it)


def serving_leg(args):
    """The request stream: `args.requests` seeded immersed-ellipse solves (grids
    64x64 .. 1024x512, Mach 1.5-4, 500 steps, IDW ghost fluid). Measures the
    plain per-solve FP64 path on a sample, the batch engine (mixed precision)
    saturated with the whole stream, the batch's latency under Poisson arrivals
    at 0.8x the plain path's rate, and the L1 density gate on a sample."""
    import time as _t
    from ibtfo_amd import serving as S
    stream = S.ellipse_stream(n=args.requests, seed=1234, steps=args.serve_steps)
    items64 = S.ellipse_items(stream, precision="FP64")
    items32 = S.ellipse_items(stream, precision="mixed")
    cells = np.array([s["N_x"] * s["N_y"] for s in stream], dtype=float)
    # plain per-solve FP64 path on the first `plain_sample` requests
    k = min(args.plain_sample, len(items64))
    _plain_solve(items64[0])  # warm-up (code objects, allocator)
    t0 = _t.perf_counter()
    for it in items64[:k]:
        _plain_solve(it)
    plain_s = _t.perf_counter() - t0
    plain_rate = k / plain_s
    plain_cu = cells[:k].sum() * args.serve_steps / plain_s
    bc = S.batch_config(Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM",
                        Precision="mixed", quirks=0, hbm_budget=args.serve_budget << 30, max_slots=args.serve_slots,
                        max_step_blocks=args.serve_step_blocks, max_ghost_blocks=4096, max_ring_blocks=8192,
                        rows=args.serve_rows, chunk=args.serve_chunk)
    _say("plain per-solve path measured")
    # strip height and launch form of the batch: the library's batch autotuner on a
    # sample of the stream resident together (probes run without a stop)
    b = S.Batch(bc)
    tuned = b.autotune(items32[: args.serve_tune_sample], probe_chunks=2)
    bc = b.bc
    b.close()
    _say("batch tile autotuned: rows %d, %s" % (tuned[0], "persistent kernel" if tuned[1] else "one launch per phase"))
    # saturation: every request queued at t = 0
    b = S.Batch(bc)
    b.submit(**items32[0])
    b.drain()  # warm-up: capture, code objects
    b.close()
    b = S.Batch(bc)
    sat = S.serve(b, items32)
    st = b.stats()
    b.close()
    sat_rate = len(items32) / sat["wall"]
    sat_cu = cells.sum() * args.serve_steps / sat["wall"]
    _say("saturated batch measured")
    # Poisson arrivals at 0.8x the plain path's solves/s
    lam = 0.8 * plain_rate
    rng = np.random.default_rng(4321)
    arr = np.cumsum(rng.exponential(1.0 / lam, size=len(items32)))
    b = S.Batch(bc)
    pois = S.serve(b, items32, arrivals=arr)
    b.close()
    lat = pois["latency"] * 1e3
    _say("Poisson-arrival batch measured")
    # L1 density gate on a sample: mixed (batched) and FP64 (plain) against a 2x-refined FP64 solve
    gate = []
    for i in range(min(args.gate_sample, len(stream))):
        b = S.Batch(bc)
        t = b.submit(**dict(items32[i], keep_state=True))
        b.drain()
        r32, U32 = b.result(t)
        b.close()
        gate.append(S.l1_gate(stream[i], items64[i], U32))
    return {
        "requests": len(items32), "steps_per_request": args.serve_steps,
        "workload": "immersed ellipse, Mach 1.5-4 inflow, grids 64x64..1024x512 (seed 1234), MUSCL + VanLeer "
                    "+ HLL, IDW ghost fluid, batch in mixed precision",
        "plain_fp64": {"solves_per_s": plain_rate, "cell_updates_per_s": plain_cu, "sample": k},
        "batched_mixed_saturated": {"solves_per_s": sat_rate, "cell_updates_per_s": sat_cu,
                                    "speedup_vs_plain_fp64": sat_rate / plain_rate, "wall_s": sat["wall"],
                                    "p50_ms": float(np.percentile(sat["latency"] * 1e3, 50)),
                                    "p99_ms": float(np.percentile(sat["latency"] * 1e3, 99)), "chunks": st.chunks},
        "batched_mixed_poisson": {"rate_per_s": lam, "p50_ms": float(np.percentile(lat, 50)),
                                  "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                                  "solves_per_s": len(items32) / pois["wall"], "wall_s": pois["wall"]},
        "l1_gate": {"rule": "L1(rho_mixed - rho_ref) <= 1.1 L1(rho_fp64 - rho_ref), rho_ref = 2x-refined FP64 "
                            "solve averaged to the grid", "samples": gate,
                    "passed": all(g["ratio"] <= 1.1 for g in gate)},
        "batch": {"slots": args.serve_slots, "budget_gib": args.serve_budget, "rows": tuned[0],
                  "launch": "persistent kernel" if tuned[1] else "one launch per phase",
                  "autotuned": {"sample": min(args.serve_tune_sample, len(items32)), "ms_per_batched_step": tuned[2]},
                  "chunk": args.serve_chunk},
    }


COPY_BYTES = 1 << 28  # calibration copy: 256 MiB read + 256 MiB written


def pmc_probe(args):
    """Workload for the two hardware-counter passes (FETCH_SIZE, then WRITE_SIZE):
    a few steps of the timed workload, then a plain device copy of known size
    (torch elementwise multiply by 1.0, COPY_BYTES read and written) whose
    counter readings calibrate the counters' units."""
    import torch
    import ibtfo_amd as ib
    from ibtfo_amd import workloads as W
    cfg = W.mach3_bodies_config(ib, N_x=args.nx, N_y=args.ny, Riemann=args.riemann, Limiter=args.limiter,
                                Boundary=args.boundary)
    s = ib.Solver(cfg)
    workload_geometry(args, s)
    s.initialize("uniform", W.MACH3_INFLOW)
    s.prepare(True)
    if args.rows > 0:
        s.set_rows(args.rows)
    s.step(args.steps)
    s.status()
    s.close()
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("pmc probe: %d steps, 3 copies of %d bytes" % (args.steps, COPY_BYTES), flush=True)
    return 0


def _pmc_means(root, counter):
    """Mean counter value per dispatch, per kernel name, from rocprofv3 CSV output."""
    import csv
    import glob
    acc = {}
    files = glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % root)
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                k = row["Kernel_Name"]
                v = acc.setdefault(k, {})
                d = row.get("Dispatch_Id", len(v))
                v[d] = v.get(d, 0.0) + float(row["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in acc.items()}


def pmc_summary(args):
    """Turns the FETCH_SIZE and WRITE_SIZE passes into HBM bytes per step-kernel
    launch: each counter is scaled by (known copy bytes) / (counter read for the
    calibration copy)."""
    fetch = _pmc_means(args.pmc_fetch, "FETCH_SIZE")
    write = _pmc_means(args.pmc_write, "WRITE_SIZE")

    def pick(d, key):
        ks = [k for k in d if key in k]
        if not ks:
            raise SystemExit("no kernel matching %r in %s" % (key, sorted(d)))
        return ks[0], d[ks[0]]

    kname, f_step = pick(fetch, "k_step")
    _, w_step = pick(write, "k_step")
    cname, f_copy = pick(fetch, "MulFunctor")
    _, w_copy = pick(write, "MulFunctor")
    sf = COPY_BYTES / (f_copy * 1024.0)
    sw = COPY_BYTES / (w_copy * 1024.0)
    cells = args.nx * args.ny
    out = dict(kernel=kname, calibration_kernel=cname, copy_bytes=COPY_BYTES,
               fetch_kib_per_launch=f_step, write_kib_per_launch=w_step,
               copy_fetch_kib=f_copy, copy_write_kib=w_copy, fetch_scale=sf, write_scale=sw,
               bytes_per_launch=f_step * 1024.0 * sf + w_step * 1024.0 * sw,
               algorithmic_bytes_per_launch=BYTES_PER_CELL * cells, cells=cells)
    out["bytes_per_cell"] = out["bytes_per_launch"] / cells
    out["workload"] = workload_name(args)
    out["tile_rows"] = args.rows or "default"
    rows = " --rows %d" % args.rows if args.rows else ""
    rs = " --riemann %s" % args.riemann if args.riemann != "HLL" else ""
    out["riemann"] = args.riemann
    out["commands"] = [
        "cd /tmp && export TMPDIR=/tmp",
        "timeout -s KILL 120 rocprofv3 --pmc FETCH_SIZE -d $OUT/pmc_fetch -o fetch --output-format csv "
        "-- python3 $REPO/bench.py --pmc-probe --steps 5" + rows + rs,
        "timeout -s KILL 120 rocprofv3 --pmc WRITE_SIZE -d $OUT/pmc_write -o write --output-format csv "
        "-- python3 $REPO/bench.py --pmc-probe --steps 5" + rows + rs,
        "python3 bench.py --pmc-fetch $OUT/pmc_fetch --pmc-write $OUT/pmc_write --pmc-out "
        "profiles/%s%s%s" % (traffic_file(args.riemann), rows, rs)]
    out["note"] = ("FETCH_SIZE / WRITE_SIZE are KiB per dispatch (mean over the step-kernel dispatches); each is "
                   "scaled by copy_bytes / (1024 * the same counter for a 256 MiB elementwise copy); $REPO is the "
                   "repository root, $OUT the folder the counter runs write to")
    print(json.dumps(out, indent=1))
    if args.pmc_out:
        with open(args.pmc_out, "w") as f:
            json.dump(out, f, indent=1)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--riemann", default="HLL")
    ap.add_argument("--limiter", default="VanLeer")
    ap.add_argument("--boundary", default="CutCell")
    ap.add_argument("--geometry", default="mask", choices=["mask", "shapes"],
                    help="bodies as a binary mask (level set on the device) or as exact signed distances")
    ap.add_argument("--no-sw", action="store_true", help="skip the Steger-Warming leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other reference configurations")
    ap.add_argument("--kernel-steps", type=int, default=500, help="steps of the per-launch kernel timing pass")
    ap.add_argument("--cpu-vortex-steps", type=int, default=24)
    ap.add_argument("--cpu-incline-steps", type=int, default=120)
    ap.add_argument("--ref-incline-steps", type=int, default=20,
                    help="steps of the inclined channel run by the reference's own build")
    ap.add_argument("--rows", type=int, default=0, help="fixed rows per strip (skips the autotuner)")
    ap.add_argument("--no-tune", action="store_true", help="default tile shape, no autotuning")
    ap.add_argument("--tune-steps", type=int, default=20)
    ap.add_argument("--graph-chunk", type=int, default=100, help="steps per captured HIP graph (0: plain launches)")
    ap.add_argument("--cpu-steps", type=int, default=40, help="steps of the CPU baseline (full grid)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps-1t", type=int, default=5, help="steps of the one-thread CPU figure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-idw", action="store_true", help="skip the IDW ghost-fluid leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-stream leg")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--gate-sample", type=int, default=4)
    ap.add_argument("--serve-budget", type=int, default=64, help="batch HBM budget in GiB")
    ap.add_argument("--serve-slots", type=int, default=256)
    ap.add_argument("--serve-step-blocks", type=int, default=65536)
    ap.add_argument("--serve-rows", type=int, default=16, help="batch strip height before the batch autotuner")
    ap.add_argument("--serve-tune-sample", type=int, default=256, help="requests the batch autotuner times")
    ap.add_argument("--serve-chunk", type=int, default=50)
    ap.add_argument("--pmc-probe", action="store_true", help="run the hardware-counter workload and exit")
    ap.add_argument("--configs-only", nargs="?", const="sod,channel,vortex", default=None,
                    help="run only these of the other configurations' legs (comma list of sod, channel, vortex; "
                         "default all) and print them (no bench line)")
    ap.add_argument("--pmc-fetch", help="rocprofv3 output folder of the FETCH_SIZE pass (summary mode)")
    ap.add_argument("--pmc-write", help="rocprofv3 output folder of the WRITE_SIZE pass (summary mode)")
    ap.add_argument("--pmc-out", help="where the summary mode writes its JSON")
    args = ap.parse_args()
    if args.pmc_fetch and args.pmc_write:
        return pmc_summary(args)
    if args.gpus != 1:
        print("bench.py: the solver keeps the whole domain on one GPU; --gpus %d is not supported "
              "(a multi-GPU slab decomposition is described, not implemented)" % args.gpus, file=sys.stderr)
        return 2
    if args.pmc_probe:
        return pmc_probe(args)
    if args.configs_only:  # the other configurations alone (profiling them), one JSON object
        print(json.dumps({"configs": config_legs(args, set(args.configs_only.split(",")))}))
        return 0
    t_start = time.perf_counter()

    def progress(what):
        print("bench: %s done (%.0f s)" % (what, time.perf_counter() - t_start), file=sys.stderr, flush=True)

    g = gpu_leg(args, keep_state=not args.no_mixed, keep_phi=not args.no_cpu)
    progress("FP64 leg")
    mixed = None
    if not args.no_mixed:
        m = gpu_leg(args, precision="mixed", keep_state=True)
        fluid = g["types"] == 0
        l1 = float(np.mean(np.abs(m["rho"][fluid] - g["rho"][fluid])))
        scale = float(np.mean(np.abs(g["rho"][fluid])))
        progress("mixed-precision leg")
        mixed = {"form": "FP32 limiter slopes and Riemann fluxes; FP64 state, face-state primitives, update, "
                         "ghost fill and reductions",
                 "value": m["value"], "unit": "cell-updates/s", "ms_per_step": m["ms_per_step"],
                 "kernel_ms": m["kernel_ms"], "speedup_vs_fp64": m["value"] / g["value"],
                 "l1_rho_vs_fp64": l1, "l1_rho_vs_fp64_rel": l1 / scale, "steps_compared": m["iter"],
                 "tile_rows": m["rows"]}
    sw = None
    if not args.no_sw and args.riemann != "StegerWarming":
        w = gpu_leg(args, riemann="StegerWarming")
        sw = {"value": w["value"], "unit": "cell-updates/s", "ms_per_step": w["ms_per_step"],
              "kernel": "k_step<MUSCL,StegerWarming,VanLeer>", "kernel_ms": w["kernel_ms"],
              "roofline_frac": w["achieved"] / HBM_PEAK_GBS, "traffic": w["traffic"], "tile_rows": w["rows"],
              "steps_taken": w["iter"], "status": w["status"]}
        progress("Steger-Warming leg")
    idw = None
    if not args.no_idw and args.boundary != "IDWBaseGFM":
        w = gpu_leg(args, boundary="IDWBaseGFM")
        idw = {"value": w["value"], "unit": "cell-updates/s", "ms_per_step": w["ms_per_step"],
               "boundary": "IDWBaseGFM (the reference's inverse-distance ghost fluid)",
               "kernel": "k_step<MUSCL,HLL,VanLeer>", "kernel_ms": w["kernel_ms"],
               "roofline_frac": w["achieved"] / HBM_PEAK_GBS, "ghost_cells": w["nghost"], "tile_rows": w["rows"],
               "steps_taken": w["iter"], "status": w["status"]}
        progress("IDW ghost-fluid leg")
    serving = None if args.no_serve else serving_leg(args)
    progress("request-stream leg")
    configs = None if args.no_configs else config_legs(args)
    progress("other configurations")
    copy_gbs = copy_bandwidth()
    cpu = None
    if not args.no_cpu:
        c = cpu_leg(args, g.pop("phi"))
        progress("CPU baseline")
        cpu = dict(value=c["value"], unit="cell-updates/s", cores=c["cores"], kind="port", sample=c["sample"],
                   ACC_NUM_CORES=c["acc_num_cores"], physical_cores=c["physical_cores"], affinity=c["affinity"],
                   single_thread={"value": c["value_1t"], "cores": 1, "sample": "the first %d steps of the same "
                                  "workload on one thread" % c["steps_1t"]})
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
        "config": {"workload": workload_name(args),
                   "N_x": args.nx, "N_y": args.ny, "ghost_cells": g["nghost"], "cut_cells": g["ncut"], "merge_groups": g["ngroups"],
                   "steps_taken": g["iter"], "status": g["status"],
                   "geometry_ms": g["geometry_ms"], "png_decode_ms": g["png_decode_ms"],
                   "tile": {"columns": STRIP_COLUMNS, "rows": g["rows"], "form": g["form"], "grid": g["grid"],
                            "autotuned": g["tuned"]},
                   "launch": ("%d timed slices of %s steps, each a captured HIP graph of %d steps replayed %s"
                              % (len(g["slices"]), "/".join(map(str, sorted(set(g["slices"])))), g["chunk"],
                                 "/".join(str(n // g["chunk"]) for n in sorted(set(g["slices"]))) + " times")
                              if g["chunk"] else "stream launches, no graph"),
                   "parallelism": "single GPU, whole domain resident"},
        "roofline": {"bound": "hbm", "achieved": g["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": g["achieved"] / HBM_PEAK_GBS, "traffic": g["traffic"],
                     "kernel": "k_step<MUSCL,HLL,VanLeer>", "kernel_ms": g["kernel_ms"],
                     "bytes_per_cell": BYTES_PER_CELL, "copy_gbs_measured": copy_gbs,
                     "frac_of_copy": g["achieved"] / copy_gbs if copy_gbs else None},
        "timing": {"slices": len(g["slice_rates"]), "slice_values_sorted": g["slice_rates"],
                   "median_slice_value": g["median"]},
        "cpu_baseline": cpu,
        "steger_warming": sw,
        "idw_ghost_fluid": idw,
        "mixed_precision": mixed,
        "serving": serving,
        "configs": configs,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
