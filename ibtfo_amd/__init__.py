"""Python front end of the MI355X-native IBTFO Euler update (ctypes over the C ABI).

The compute path is libibtfo_amd.so (hand-written gfx950 HIP kernels behind
include/ibtfo.h). This module only marshals arguments; there is no Python or
CPU fallback: if the library is missing or no GPU is present, calls fail.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibtfo_amd.so")
RUN_PATH = os.path.join(HERE, "ibtfo_run")

SCHEME = {"Lax": 0, "MUSCL": 1, "MHM": 2, "Godunov": 3}
RIEMANN = {"HLL": 0, "HLLC": 1, "PVRS": 2, "StegerWarming": 3}
LIMITER = {"minmod": 0, "VanLeer": 1, "VanAlbada": 2}
BOUNDARY = {"Regular": 0, "IDWBaseGFM": 1, "LinearBaseGFM": 2, "RiemannBaseGFM": 3, "BCGFM": 4, "HBCGFM": 5,
            "CutCell": 6, "SimpleBound": 7}
TIME = {"SingleStep": 0, "RK2": 1}
EDGE = {"reference": 0, "periodic": 1, "inflow": 2, "pnggeo": 3}
PRECISION = {"FP64": 0, "fp64": 0, "mixed": 1}
LEVELSET = {"none": 0, "channel": 1, "annulus": 2, "field": 3, "ellipse": 4, "shapes": 5}
INIT = {"sod_rows": 0, "incline": 1, "vortex": 2, "sod_x": 3, "uniform": 4, "shu": 5, "shock": 6}

OK, E_ARG, E_HIP, E_NOMEM, E_IO, E_STATE, E_NONFINITE = range(7)


class Config(C.Structure):
    _fields_ = [("N_x", C.c_int), ("N_y", C.c_int), ("num_ghost_cell", C.c_int), ("max_iter", C.c_int),
                ("lo_x", C.c_double), ("lo_y", C.c_double), ("hi_x", C.c_double), ("hi_y", C.c_double),
                ("Psy_time", C.c_double), ("CFL_number", C.c_double), ("gamma", C.c_double),
                ("scheme", C.c_int), ("riemann", C.c_int), ("limiter", C.c_int), ("boundary", C.c_int),
                ("time", C.c_int), ("edge", C.c_int), ("precision", C.c_int), ("quirks", C.c_int),
                ("edge_state", C.c_double * 4)]


class IbtfoError(RuntimeError):
    def __init__(self, rc, what, msg):
        super().__init__("%s failed (%d): %s" % (what, rc, msg))
        self.rc = rc


_lib = None
_D = C.POINTER(C.c_double)


def lib():
    """Load libibtfo_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libibtfo_amd.so is not built; run __graft_entry__.build() or make -C ibtfo_amd/csrc")
        L = C.CDLL(LIB_PATH)
        L.ibtfo_last_error.restype = C.c_char_p
        L.ibtfo_footprint.restype = C.c_longlong
        L.ibtfo_footprint.argtypes = [C.POINTER(Config)]
        L.ibtfo_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        for name in ["ibtfo_destroy", "ibtfo_step", "ibtfo_step_graph", "ibtfo_status", "ibtfo_log",
                     "ibtfo_compute_dt", "ibtfo_mass", "ibtfo_boundary", "ibtfo_prepare", "ibtfo_geometry",
                     "ibtfo_initialize", "ibtfo_upload_state", "ibtfo_download_state", "ibtfo_download_geometry",
                     "ibtfo_download_scheme_index", "ibtfo_sch_width", "ibtfo_num_ghosts", "ibtfo_dims",
                     "ibtfo_time_steps", "ibtfo_device_state", "ibtfo_set_stream", "ibtfo_set_rows",
                     "ibtfo_set_tile", "ibtfo_tile_ex", "ibtfo_one_kernel_steps", "ibtfo_tile", "ibtfo_autotune", "ibtfo_geometry_mask", "ibtfo_geometry_png",
                     "ibtfo_png_decode", "ibtfo_cut_counts", "ibtfo_write_vtk", "ibtfo_vtk_write_host"]:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _check(rc, what):
    if rc != OK:
        raise IbtfoError(rc, what, lib().ibtfo_last_error().decode(errors="replace"))


def config(**kw):
    """Config with the reference driver defaults, overridden by keyword (slot
    names may be given as strings, e.g. Scheme='MHM')."""
    c = Config()
    lib().ibtfo_config_default(C.byref(c))
    names = {"Scheme": ("scheme", SCHEME), "Riemann": ("riemann", RIEMANN), "Limiter": ("limiter", LIMITER),
             "Boundary": ("boundary", BOUNDARY), "Time": ("time", TIME), "Edge": ("edge", EDGE),
             "Precision": ("precision", PRECISION)}
    for k, v in kw.items():
        if k in names:
            field, table = names[k]
            setattr(c, field, table[v] if isinstance(v, str) else int(v))
        elif k == "edge_state":
            for q in range(4):
                c.edge_state[q] = float(v[q])
        else:
            setattr(c, k, v)
    return c



def write_vtk_host(cfg, iter_, now_t, U_OLD, XYCOORD, directory):
    """WriteData's files for host arrays in the reference's AoS layout (no GPU needed)."""
    U = np.ascontiguousarray(U_OLD, dtype=np.float64)
    XY = np.ascontiguousarray(XYCOORD, dtype=np.float64)
    _check(lib().ibtfo_vtk_write_host(C.byref(cfg), int(iter_), C.c_double(now_t), ptr(U), ptr(XY),
                                      str(directory).encode()), "ibtfo_vtk_write_host")


def png_decode(path):
    """(H, W, 4) uint8 RGBA of a PNG file as the mask level set sees it."""
    w, h = C.c_int(), C.c_int()
    with open(path, "rb") as f:
        head = f.read(24)
    hw = int.from_bytes(head[16:20], "big") if len(head) == 24 else 0
    hh = int.from_bytes(head[20:24], "big") if len(head) == 24 else 0
    if head[:8] == b"\x89PNG\r\n\x1a\n" and head[12:16] == b"IHDR" and 0 < hw <= 65536 and 0 < hh <= 65536:
        w.value, h.value = hw, hh  # the size from the header; the decode below checks the rest
    else:  # let the decoder say what is wrong
        _check(lib().ibtfo_png_decode(os.fsencode(path), None, C.c_longlong(0), C.byref(w), C.byref(h)),
               "ibtfo_png_decode")
    out = np.zeros((h.value, w.value, 4), dtype=np.uint8)
    _check(lib().ibtfo_png_decode(os.fsencode(path), out.ctypes.data_as(C.c_void_p), C.c_longlong(out.size),
                                  C.byref(w), C.byref(h)), "ibtfo_png_decode")
    return out


RECORD_W = 28  # doubles per cell of the scheme's scratch record (7 slots x 4)


def flux_record(gamma, rec, u_slot, f_slot=None, g_slot=None):
    """Conserve2Flux{,X,Y} on the GPU for every cell of a host scratch record
    (M*N*28 doubles): the state in slot u_slot, its x flux to slot f_slot and
    its y flux to slot g_slot (in place)."""
    assert rec.dtype == np.float64 and rec.flags.c_contiguous and rec.size % RECORD_W == 0
    base = rec.ctypes.data

    def at(slot):
        return None if slot is None else C.cast(base + 32 * slot, _D)
    _check(lib().ibtfo_flux_record_host(C.c_double(gamma), at(u_slot), at(f_slot), at(g_slot),
                                        C.c_longlong(rec.size // RECORD_W), RECORD_W), "ibtfo_flux_record_host")


def flux_record_device(gamma, u_ptr, f_ptr, g_ptr, ncells, stride, stream=None):
    """ibtfo_flux_record on device addresses (0 or None for an absent f or g)."""
    _check(lib().ibtfo_flux_record(C.c_double(gamma), C.c_void_p(u_ptr), C.c_void_p(f_ptr or None),
                                   C.c_void_p(g_ptr or None), C.c_longlong(ncells), int(stride),
                                   C.c_void_p(stream)), "ibtfo_flux_record")


def config_from_inp(path):
    c = Config()
    _check(lib().ibtfo_config_from_inp(path.encode(), C.byref(c)), "ibtfo_config_from_inp")
    return c


class Solver:
    """One solve on the GPU (a wrapper of ibtfo_ctx)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.h = C.c_void_p()
        _check(lib().ibtfo_create(C.byref(cfg), C.byref(self.h)), "ibtfo_create")
        M, N = C.c_int(), C.c_int()
        lib().ibtfo_dims(self.h, C.byref(M), C.byref(N))
        self.M, self.N = M.value, N.value

    def close(self):
        if self.h:
            lib().ibtfo_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry_mask(self, rgba):
        """Mask level set from an (H, W, 4) uint8 image, top row first."""
        img = np.ascontiguousarray(rgba, dtype=np.uint8)
        assert img.ndim == 3 and img.shape[2] == 4
        _check(lib().ibtfo_geometry_mask(self.h, img.ctypes.data_as(C.c_void_p), img.shape[1], img.shape[0]),
               "ibtfo_geometry_mask")

    def geometry_png(self, path):
        _check(lib().ibtfo_geometry_png(self.h, os.fsencode(path)), "ibtfo_geometry_png")

    def geometry(self, kind="none", prm=(), phi=None):
        k = LEVELSET[kind] if isinstance(kind, str) else int(kind)
        p = np.ascontiguousarray(prm, dtype=np.float64)
        ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        _check(lib().ibtfo_geometry(self.h, k, ptr(p) if p.size else None, int(p.size), ptr(ph)), "ibtfo_geometry")

    def initialize(self, kind, prm):
        k = INIT[kind] if isinstance(kind, str) else int(kind)
        p = np.ascontiguousarray(prm, dtype=np.float64)
        _check(lib().ibtfo_initialize(self.h, k, ptr(p), int(p.size)), "ibtfo_initialize")

    def upload(self, U_OLD, U_NEW=None):
        _check(lib().ibtfo_upload_state(self.h, ptr(np.ascontiguousarray(U_OLD)),
                                        ptr(None if U_NEW is None else np.ascontiguousarray(U_NEW))),
               "ibtfo_upload_state")

    def prepare(self, apply_boundary):
        _check(lib().ibtfo_prepare(self.h, int(apply_boundary)), "ibtfo_prepare")

    def step(self, n, graph_chunk=0):
        if graph_chunk:
            _check(lib().ibtfo_step_graph(self.h, int(n), int(graph_chunk)), "ibtfo_step_graph")
        else:
            _check(lib().ibtfo_step(self.h, int(n)), "ibtfo_step")

    def run(self, max_steps=None, graph_chunk=100, block=2000):
        """Steps in blocks until the device stops the run (Psy_time or max_iter)
        or `max_steps` more steps were taken; returns the steps taken. Blocks keep
        the host from queueing steps the device would skip."""
        start = done = self.status()[0]
        while max_steps is None or done - start < max_steps:
            n = block if max_steps is None else min(block, max_steps - (done - start))
            self.step(n, graph_chunk=graph_chunk if n >= graph_chunk else 0)
            it = self.status()[0]
            if it < done + n:
                return it - start
            done = it
        return done - start

    def status(self):
        it, t, dt = C.c_int(), C.c_double(), C.c_double()
        rc = lib().ibtfo_status(self.h, C.byref(it), C.byref(t), C.byref(dt))
        if rc not in (OK, E_NONFINITE):
            _check(rc, "ibtfo_status")
        return it.value, t.value, dt.value, rc

    def log(self, count):
        d, t, m = np.zeros(count), np.zeros(count), np.zeros(count)
        if count:
            _check(lib().ibtfo_log(self.h, int(count), ptr(d), ptr(t), ptr(m)), "ibtfo_log")
        return d, t, m

    def state(self):
        U = np.zeros(self.M * self.N * 4)
        _check(lib().ibtfo_download_state(self.h, ptr(U)), "ibtfo_download_state")
        return U

    def xycoord(self):
        XY = np.zeros(self.M * self.N * 6)
        _check(lib().ibtfo_download_geometry(self.h, ptr(XY)), "ibtfo_download_geometry")
        return XY

    def scheme_index(self):
        W = lib().ibtfo_sch_width(self.h)
        S = np.zeros(self.M * self.N * W)
        _check(lib().ibtfo_download_scheme_index(self.h, ptr(S)), "ibtfo_download_scheme_index")
        return S

    def num_ghosts(self):
        return lib().ibtfo_num_ghosts(self.h)

    def write_vtk(self, directory, iter_, now_t):
        """WriteData's files (dat_0_<iter>.vtk, result.visit appended) for the current state."""
        _check(lib().ibtfo_write_vtk(self.h, str(directory).encode(), int(iter_), C.c_double(now_t)),
               "ibtfo_write_vtk")

    def cut_counts(self):
        """(cut cells, small-cell merge groups) of the cut-cell boundary."""
        a, b = C.c_int(), C.c_int()
        _check(lib().ibtfo_cut_counts(self.h, C.byref(a), C.byref(b)), "ibtfo_cut_counts")
        return a.value, b.value

    def compute_dt(self):
        v = C.c_double()
        _check(lib().ibtfo_compute_dt(self.h, C.byref(v)), "ibtfo_compute_dt")
        return v.value

    def mass(self):
        v = C.c_double()
        _check(lib().ibtfo_mass(self.h, C.byref(v)), "ibtfo_mass")
        return v.value

    def boundary(self):
        _check(lib().ibtfo_boundary(self.h), "ibtfo_boundary")

    def riemann_record(self, rec, ydir):
        """Riemann_solverX (ydir 0) / Riemann_solverY (1) with this solve's
        Riemann slot, gamma and quirks on a host scratch record (M*N*28
        doubles, updated in place)."""
        assert rec.dtype == np.float64 and rec.flags.c_contiguous and rec.size == self.M * self.N * RECORD_W
        _check(lib().ibtfo_riemann_record_host(self.h, int(ydir), ptr(rec)), "ibtfo_riemann_record_host")

    def riemann_record_device(self, rec_ptr, ydir):
        """The same on a device record (an address, e.g. a tensor's data_ptr());
        queued on the solve's stream."""
        _check(lib().ibtfo_riemann_record(self.h, int(ydir), C.c_void_p(rec_ptr)), "ibtfo_riemann_record")

    def set_rows(self, rows):
        _check(lib().ibtfo_set_rows(self.h, int(rows)), "ibtfo_set_rows")

    def set_tile(self, rows, row_parallel=False):
        """rows per strip and the form (row_parallel: one wave per row, 1..7 rows)."""
        _check(lib().ibtfo_set_tile(self.h, int(rows), int(bool(row_parallel))), "ibtfo_set_tile")

    def tile(self):
        r, a, b = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ibtfo_tile(self.h, C.byref(r), C.byref(a), C.byref(b)), "ibtfo_tile")
        return r.value, a.value, b.value

    def one_kernel_steps(self):
        """True when a step runs as one kernel (the deferred finish of small grids)."""
        return lib().ibtfo_one_kernel_steps(self.h) == 1

    def tile_form(self):
        """True when the row-parallel step kernel is selected."""
        p = C.c_int()
        _check(lib().ibtfo_tile_ex(self.h, None, None, None, C.byref(p)), "ibtfo_tile_ex")
        return bool(p.value)

    def autotune(self, probe_steps=20):
        """Picks the step kernel's rows per strip by timing; returns (rows, ms/step)."""
        r, ms = C.c_int(), C.c_double()
        _check(lib().ibtfo_autotune(self.h, int(probe_steps), C.byref(r), C.byref(ms)), "ibtfo_autotune")
        return r.value, ms.value

    def time_steps(self, n, kernel=True):
        tot, ker = C.c_double(), C.c_double()
        _check(lib().ibtfo_time_steps(self.h, int(n), C.byref(tot), C.byref(ker) if kernel else None),
               "ibtfo_time_steps")
        return tot.value, (ker.value if kernel else None)


def run_case(cfg, levelset=("none", ()), init=("sod_rows", ()), nsteps=None, apply_boundary=None, graph_chunk=0,
             chunk=256):
    """Reference driver order on the GPU: geometry, Initialize, [Boundary],
    ComputeDt, then steps until the device stops (t_end / max_iter) or `nsteps`."""
    s = Solver(cfg)
    s.geometry(levelset[0], levelset[1])
    s.initialize(init[0], init[1])
    if apply_boundary is None:
        apply_boundary = cfg.boundary != BOUNDARY["Regular"]
    s.prepare(apply_boundary)
    limit = cfg.max_iter if nsteps is None else min(nsteps, cfg.max_iter)
    done = 0
    while done < limit:
        n = min(chunk, limit - done)
        s.step(n, graph_chunk)
        it, _, _, _ = s.status()
        if it - done < n:
            done = it
            break
        done = it
    return s
