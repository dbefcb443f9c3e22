"""Request serving on the batch engine (include/ibtfo.h, "Batch engine").

`Batch` wraps ibtfo_batch: requests of different grid sizes and geometries are
admitted by HBM budget into one device-resident, graph-captured batched loop.
`serve` drives a stream of timed arrivals through it and reports per-request
latency (arrival -> result on the host) and throughput. `ellipse_stream` makes
the seeded synthetic stream of immersed-ellipse solves.
"""
import ctypes as C
import math
import time

import numpy as np

from . import (BOUNDARY, E_NONFINITE, LIMITER, OK, PRECISION, RIEMANN, SCHEME, TIME, Config, _check, config, lib)

_D = C.POINTER(C.c_double)


class BatchConfig(C.Structure):
    _fields_ = [("scheme", C.c_int), ("riemann", C.c_int), ("limiter", C.c_int), ("boundary", C.c_int),
                ("time", C.c_int), ("precision", C.c_int), ("quirks", C.c_int), ("hbm_budget", C.c_longlong),
                ("max_slots", C.c_int), ("max_step_blocks", C.c_int), ("max_ghost_blocks", C.c_int),
                ("max_ring_blocks", C.c_int), ("rows", C.c_int), ("chunk", C.c_int), ("persistent", C.c_int)]


class Request(C.Structure):
    _fields_ = [("cfg", Config), ("ls_kind", C.c_int), ("ls_nprm", C.c_int), ("ls_prm", _D), ("phi", _D),
                ("init_kind", C.c_int), ("init_nprm", C.c_int), ("init_prm", _D), ("keep_state", C.c_int),
                ("arrival", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iter", C.c_int), ("now_t", C.c_double), ("dt", C.c_double),
                ("mass", C.c_double), ("M", C.c_int), ("N", C.c_int), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("active", C.c_int), ("queued", C.c_int), ("finished", C.c_int), ("chunks", C.c_longlong),
                ("steps", C.c_longlong), ("cell_updates", C.c_longlong), ("arena_used", C.c_longlong),
                ("arena_capacity", C.c_longlong), ("busy_s", C.c_double)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    L.ibtfo_batch_clock.restype = C.c_double
    L.ibtfo_batch_clock.argtypes = [C.c_void_p]
    for n in ["ibtfo_batch_config_default", "ibtfo_batch_create", "ibtfo_batch_destroy", "ibtfo_batch_submit",
              "ibtfo_batch_progress", "ibtfo_batch_poll", "ibtfo_batch_result", "ibtfo_batch_release",
              "ibtfo_batch_stats_get", "ibtfo_batch_autotune"]:
        getattr(L, n).restype = C.c_int
    _bound = True


def batch_config(**kw):
    """Batch settings: the library defaults overridden by keyword (slot names may
    be given as strings, e.g. Scheme='MUSCL', Precision='mixed')."""
    _bind()
    bc = BatchConfig()
    _check(lib().ibtfo_batch_config_default(C.byref(bc)), "ibtfo_batch_config_default")
    names = {"Scheme": ("scheme", SCHEME), "Riemann": ("riemann", RIEMANN), "Limiter": ("limiter", LIMITER),
             "Boundary": ("boundary", BOUNDARY), "Time": ("time", TIME), "Precision": ("precision", PRECISION)}
    for k, v in kw.items():
        if k in names:
            f, table = names[k]
            setattr(bc, f, table[v] if isinstance(v, str) else int(v))
        else:
            setattr(bc, k, v)
    return bc


def _arr(x):
    if x is None or len(x) == 0:
        return None, 0
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a, len(a)


class Batch:
    """One batch engine (ibtfo_batch) on the current GPU."""

    def __init__(self, bc):
        _bind()
        self.bc = bc
        self.h = C.c_void_p()
        _check(lib().ibtfo_batch_create(C.byref(bc), C.byref(self.h)), "ibtfo_batch_create")
        self._keep = {}

    def close(self):
        if self.h:
            lib().ibtfo_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return lib().ibtfo_batch_clock(self.h)

    def autotune(self, items, probe_chunks=2):
        """Picks the batch's strip height and launch mode (persistent kernel or
        one launch per phase) by timing the sample `items` (submit() keyword
        dicts) resident together; the batch must be idle. Returns (rows,
        persistent, ms per batched step)."""
        reqs, keep = [], []
        for it in items:
            r, k = self._request(**it)
            reqs.append(r)
            keep.append(k)
        arr = (Request * len(reqs))(*reqs)
        rows, per, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().ibtfo_batch_autotune(self.h, arr, len(reqs), int(probe_chunks), C.byref(rows), C.byref(per),
                                          C.byref(ms)), "ibtfo_batch_autotune")
        self.bc.rows, self.bc.persistent = rows.value, per.value
        return rows.value, per.value, ms.value

    def _request(self, cfg, levelset=("none", ()), init=("uniform", (1.0, 0.0, 0.0, 1.0)), keep_state=False,
                 arrival=-1.0, phi=None):
        """A Request struct and the arrays it points to (kept alive by the caller)."""
        from . import INIT, LEVELSET
        r = Request()
        r.cfg = cfg
        ls_kind, ls_prm = levelset
        r.ls_kind = LEVELSET[ls_kind] if isinstance(ls_kind, str) else int(ls_kind)
        lp, r.ls_nprm = _arr(ls_prm)
        r.ls_prm = lp.ctypes.data_as(_D) if lp is not None else None
        ph = None
        if phi is not None:
            ph = np.ascontiguousarray(np.asarray(phi, dtype=np.float64))
            r.phi = ph.ctypes.data_as(_D)
        ik, iprm = init
        r.init_kind = INIT[ik] if isinstance(ik, str) else int(ik)
        ip, r.init_nprm = _arr(iprm)
        r.init_prm = ip.ctypes.data_as(_D) if ip is not None else None
        r.keep_state = int(bool(keep_state))
        r.arrival = float(arrival)
        return r, (lp, ph, ip)

    def submit(self, cfg, levelset=("none", ()), init=("uniform", (1.0, 0.0, 0.0, 1.0)), keep_state=False,
               arrival=-1.0, phi=None):
        from . import INIT, LEVELSET
        r = Request()
        r.cfg = cfg
        ls_kind, ls_prm = levelset
        r.ls_kind = LEVELSET[ls_kind] if isinstance(ls_kind, str) else int(ls_kind)
        lp, r.ls_nprm = _arr(ls_prm)
        r.ls_prm = lp.ctypes.data_as(_D) if lp is not None else None
        ph = None
        if phi is not None:
            ph = np.ascontiguousarray(np.asarray(phi, dtype=np.float64))
            r.phi = ph.ctypes.data_as(_D)
        ik, iprm = init
        r.init_kind = INIT[ik] if isinstance(ik, str) else int(ik)
        ip, r.init_nprm = _arr(iprm)
        r.init_prm = ip.ctypes.data_as(_D) if ip is not None else None
        r.keep_state = int(bool(keep_state))
        r.arrival = float(arrival)
        t = C.c_longlong()
        _check(lib().ibtfo_batch_submit(self.h, C.byref(r), C.byref(t)), "ibtfo_batch_submit")
        self._keep[t.value] = bool(keep_state)
        return t.value

    def progress(self):
        a, q, f = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ibtfo_batch_progress(self.h, C.byref(a), C.byref(q), C.byref(f)), "ibtfo_batch_progress")
        return a.value, q.value, f.value

    def poll(self):
        out = []
        t = C.c_longlong()
        while lib().ibtfo_batch_poll(self.h, C.byref(t)) == 1:
            out.append(t.value)
        return out

    def result(self, ticket, state=None):
        """(Result, U) of a retired request; U (AoS, M*N*4) if it kept its state."""
        r = Result()
        _check(lib().ibtfo_batch_result(self.h, C.c_longlong(ticket), C.byref(r), None), "ibtfo_batch_result")
        U = None
        if self._keep.get(ticket, False) if state is None else state:
            U = np.zeros(r.M * r.N * 4)
            _check(lib().ibtfo_batch_result(self.h, C.c_longlong(ticket), None, U.ctypes.data_as(_D)),
                   "ibtfo_batch_result")
        return r, U

    def release(self, ticket):
        lib().ibtfo_batch_release(self.h, C.c_longlong(ticket))
        self._keep.pop(ticket, None)

    def stats(self):
        s = BatchStats()
        _check(lib().ibtfo_batch_stats_get(self.h, C.byref(s)), "ibtfo_batch_stats_get")
        return s

    def drain(self, max_rounds=10 ** 7):
        """Runs progress rounds until nothing is queued or active; returns tickets in finish order."""
        done = []
        for _ in range(max_rounds):
            a, q, _ = self.progress()
            done += self.poll()
            if a == 0 and q == 0:
                break
        return done


def serve(batch, items, arrivals=None, on_result=None):
    """Drives timed arrivals through a batch. items: list of submit() keyword
    dicts; arrivals: offsets in seconds from the start (None: all at once).
    Returns dict(ticket -> Result, 'order': tickets by completion, 'wall': s,
    'latency': array of t_done - t_arrival in s)."""
    n = len(items)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=float)
    order = np.argsort(arrivals, kind="stable")
    t0 = batch.clock()
    k = 0
    results = {}
    finished = []
    tickets = {}
    while len(finished) < n:
        now = batch.clock() - t0
        while k < n and arrivals[order[k]] <= now:
            i = order[k]
            tickets[batch.submit(arrival=t0 + arrivals[i], **items[i])] = i
            k += 1
        a, q, _ = batch.progress()
        for t in batch.poll():
            r, U = batch.result(t)
            results[tickets[t]] = r
            finished.append(tickets[t])
            if on_result is not None:
                on_result(tickets[t], r, U)
            batch.release(t)
        if a == 0 and q == 0 and k < n:  # idle until the next arrival
            wait = arrivals[order[k]] - (batch.clock() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.05))
    wall = batch.clock() - t0
    lat = np.array([results[i].t_done - results[i].t_arrival for i in range(n)])
    return dict(results=results, order=finished, wall=wall, latency=lat)


def ellipse_stream(n=512, seed=1234, steps=500, cfl=0.5, gamma=1.4):
    """Seeded immersed-ellipse solves: grid N_x = 64*16^u, N_y = 64*8^u (u ~ U(0,1),
    rounded to multiples of 8; 64x64 .. 1024x512), domain [0, N_x/N_y] x [0, 1]
    (dx = dy), an ellipse centred at (0.3 L, 0.5 H) with semi-major axis
    U(0.05, 0.15) H, axis ratio U(1, 4) and angle U(0, 180) degrees, Mach
    U(1.5, 4) inflow with rho = 1.4, p = 1 (sound speed 1), `steps` steps.
    Returns a list of dicts: N_x, N_y, L, ellipse (xc, yc, a, b, angle), inflow."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        u = rng.uniform()
        nx = int(round(64 * 16 ** u / 8)) * 8
        ny = int(round(64 * 8 ** u / 8)) * 8
        L = nx / ny
        a = rng.uniform(0.05, 0.15)
        ratio = rng.uniform(1.0, 4.0)
        ang = rng.uniform(0.0, 180.0)
        mach = rng.uniform(1.5, 4.0)
        rho, p = 1.4, 1.0
        c = math.sqrt(gamma * p / rho)
        out.append(dict(N_x=nx, N_y=ny, L=L, ellipse=(0.3 * L, 0.5, a, a / ratio, ang),
                        inflow=(rho, mach * c, 0.0, p), steps=steps, cfl=cfl))
    return out


def ellipse_items(stream, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM",
                  Time="SingleStep", precision="mixed", quirks=0, keep_state=False):
    """submit() keyword dicts for an ellipse_stream."""
    items = []
    for s in stream:
        cfg = config(N_x=s["N_x"], N_y=s["N_y"], lo_x=0.0, lo_y=0.0, hi_x=s["L"], hi_y=1.0, Psy_time=1e30,
                     max_iter=s["steps"], CFL_number=s["cfl"], gamma=1.4, Scheme=Scheme, Riemann=Riemann,
                     Limiter=Limiter, Boundary=Boundary, Time=Time, Edge="inflow", quirks=quirks,
                     edge_state=s["inflow"])
        cfg.precision = PRECISION[precision]
        items.append(dict(cfg=cfg, levelset=("ellipse", s["ellipse"]), init=("uniform", s["inflow"]),
                          keep_state=keep_state))
    return items


def plain_solve(item, chunk=50):
    """One request through the single-solve path (create, geometry, initial
    state, boundary + dt, graph-replayed steps, final status). Returns (status
    tuple, U or None, cell types or None); state and types when the item keeps
    its state."""
    from . import Solver
    s = Solver(item["cfg"])
    try:
        s.geometry(*item["levelset"])
        s.initialize(*item["init"])
        s.prepare(True)
        n, done = item["cfg"].max_iter, 0
        while done < n:  # stops early when the solve reaches its Psy_time
            m = min(chunk, n - done)
            s.step(m, graph_chunk=chunk if m == chunk else 0)
            i = s.status()[0]
            if i < done + m:
                break
            done = i
        st = s.status()
        U = s.state() if item.get("keep_state") else None
        types = s.xycoord()[5::6] if item.get("keep_state") else None
    finally:
        s.close()
    return st, U, types


def l1_gate(entry, item64, U32, fp64=None):
    """The fixed-L1 gate of one request: L1(rho_mixed - rho_ref) against
    L1(rho_fp64 - rho_ref) over the fluid cells, rho_ref being an FP64 solve on
    the 2x-refined grid to the same time, averaged to the request's grid (its
    four children all fluid). U32: the mixed-precision final state (AoS);
    fp64: (status, U, types) of the FP64 solve if already run. Returns a dict."""
    st64, U64, types = fp64 if fp64 is not None else plain_solve(dict(item64, keep_state=True))
    fine = dict(entry, N_x=2 * entry["N_x"], N_y=2 * entry["N_y"], steps=10 ** 9)
    itf = ellipse_items([fine], precision="FP64", keep_state=True)[0]
    itf["cfg"].Psy_time = st64[1]
    _, Uf, tf = plain_solve(itf)
    nx, ny = entry["N_x"], entry["N_y"]
    M, N = nx + 4, ny + 4
    rf = Uf[0::4].reshape(2 * ny + 4, 2 * nx + 4)[2:-2, 2:-2]
    tff = tf.reshape(2 * ny + 4, 2 * nx + 4)[2:-2, 2:-2]
    ref = 0.25 * (rf[0::2, 0::2] + rf[1::2, 0::2] + rf[0::2, 1::2] + rf[1::2, 1::2])
    okf = (tff[0::2, 0::2] == 0) & (tff[1::2, 0::2] == 0) & (tff[0::2, 1::2] == 0) & (tff[1::2, 1::2] == 0)
    mask = (types.reshape(N, M)[2:-2, 2:-2] == 0) & okf
    r64 = U64[0::4].reshape(N, M)[2:-2, 2:-2]
    r32 = U32[0::4].reshape(N, M)[2:-2, 2:-2]
    e64 = float(np.mean(np.abs(r64[mask] - ref[mask])))
    e32 = float(np.mean(np.abs(r32[mask] - ref[mask])))
    return dict(grid=[nx, ny], l1_fp64=e64, l1_mixed=e32, ratio=e32 / e64)


__all__ = ["Batch", "BatchConfig", "Request", "Result", "batch_config", "serve", "ellipse_stream",
           "ellipse_items", "plain_solve", "l1_gate", "OK", "E_NONFINITE"]
