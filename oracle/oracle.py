"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the result to compare against. The product (ibtfo_amd/) never imports this
module.

`Case` describes one problem the way the reference's .inp file + Makefile slots
do; `run_oracle` drives liboracle.so (the C++ restatement) in the order of
Test/Sod/Sod.cpp:94-181 / Test/InclineSod/InclineSod.cpp:73-165, and `run_ref`
drives a reference library built by build_ref.py from the reference's sources.
"""
import ctypes as C
import os
import tempfile
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

SCHEMES = {"Lax": 0, "MUSCL": 1, "MHM": 2, "Godunov": 3}
RIEMANNS = {"HLL": 0, "HLLC": 1, "PVRS": 2, "StegerWarming": 3}
LIMITERS = {"minmod": 0, "VanLeer": 1, "VanAlbada": 2}
BOUNDS = {"Regular": 0, "IDWBaseGFM": 1, "LinearBaseGFM": 2, "RiemannBaseGFM": 3, "BCGFM": 4, "HBCGFM": 5,
          "CutCell": 6, "SimpleBound": 7}
EDGES = {"reference": 0, "periodic": 1, "inflow": 2, "pnggeo": 3}
TIMES = {"SingleStep": 0, "RK2": 1}
# level-set hooks and initialisers (see or_level_set_hook / or_initialize)
LS = {"sod": 0, "sodx": 0, "incline": 1, "vortex": 2, "phi": 3, "none": 0, "pngshock": 1}
INIT = {"sod": 0, "incline": 1, "vortex": 2, "sodx": 3, "uniform": 4, "shu": 5, "pngshock": 6}

_D = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = C.CDLL(LIB)
        L.or_compute_dt.restype = C.c_double
        L.or_mass.restype = C.c_double
        L.or_limiter.restype = C.c_double
        L.or_limiter.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.or_run.restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


@dataclass
class Case:
    """A problem in the reference's terms: .inp numbers + slot selection + hooks."""
    N_x: int = 40
    N_y: int = 10
    lo_x: float = 0.0
    lo_y: float = 0.0
    hi_x: float = 1.0
    hi_y: float = 0.25
    Psy_time: float = 0.2
    max_iter: int = 100000
    CFL_number: float = 0.5
    num_ghost_cell: int = 2
    gamma: float = 1.4
    Scheme: str = "MUSCL"
    Riemann: str = "HLL"
    Limiter: str = "minmod"
    Boundary: str = "Regular"
    Time: str = "SingleStep"
    case: str = "sodx"          # hook set: sod | sodx | incline | vortex | uniform | shu
    edge: str = "reference"     # Global_Boundary policy
    quirks: int = 1             # 1 = reference arithmetic incl. its defects, 0 = corrected
    extra: dict = field(default_factory=dict)  # problem keys (rho_L.., angle, AxisD, fa, f0, ...)
    phi: object = None          # per-cell level set when case uses a given field

    @property
    def L(self):
        return self.hi_x - self.lo_x

    @property
    def H(self):
        return self.hi_y - self.lo_y

    @property
    def M(self):
        return self.N_x + 2 * self.num_ghost_cell

    @property
    def N(self):
        return self.N_y + 2 * self.num_ghost_cell

    def inp_text(self):
        keys = dict(max_iter=self.max_iter, N_x=self.N_x, N_y=self.N_y, Psy_time=self.Psy_time,
                    lo_x=self.lo_x, lo_y=self.lo_y, hi_x=self.hi_x, hi_y=self.hi_y,
                    CFL_number=self.CFL_number, num_ghost_cell=self.num_ghost_cell, gamma=self.gamma)
        keys.update(self.extra)
        lines = ["#generated"] + ["%s=%s" % (k, _num(v)) for k, v in keys.items()]
        lines += ["Scheme=%s" % self.Scheme, "Riemann=%s" % self.Riemann, "Limiter=%s" % self.Limiter,
                  "Boundary=%s" % self.Boundary]
        return "\n".join(lines) + "\n"

    def ls_prm(self):
        e = self.extra
        if self.case in ("incline", "pngshock"):
            return [e.get("angle", 30.0), e.get("AxisD", 0.4)]
        if self.case == "vortex":
            return [e.get("f0", 0.0), e.get("f1", 1.0)]
        return [0.0]

    def init_prm(self):
        e = self.extra
        left = [e.get("rho_L", 1), e.get("u_L", 0), e.get("v_L", 0), e.get("p_L", 1)]
        right = [e.get("rho_R", 0.125), e.get("u_R", 0), e.get("v_R", 0), e.get("p_R", 0.1)]
        if self.case == "sod":
            return left + right
        if self.case == "incline":
            return [e.get("angle", 30.0)] + left + right
        if self.case == "vortex":
            return [e.get("fa", 0.0), e.get("fw", 0.0)]
        if self.case == "sodx":
            return [e.get("x0", 0.5)] + left + right
        if self.case == "uniform":
            return [e.get("rho", 1.0), e.get("u", 0.0), e.get("v", 0.0), e.get("p", 1.0)]
        if self.case == "shu":
            return [e.get("x0", 5.0), e.get("y0", 5.0), e.get("beta", 5.0), 1.0, e.get("u", 1.0),
                    e.get("v", 1.0), 1.0]
        if self.case == "pngshock":  # Test/PNGGeo/Initialize.cpp: shock at x = 0.1, Mach 1.3
            return [0.1, 1.3] + left
        raise ValueError(self.case)

    def edge_prm(self):
        e = self.extra
        if self.edge == "pnggeo" and "rho" not in e:
            return pnggeo_inflow(self.gamma)
        return [e.get("rho", 1.0), e.get("u", 0.0), e.get("v", 0.0), e.get("p", 1.0)]

    def init_boundary(self):
        # Sod.cpp applies no Boundary before the loop; InclineSod.cpp:81 and Vortex.cpp do
        return 0 if self.Boundary == "Regular" else 1


def pnggeo_inflow(gamma, rho1=1.0, p1=0.71429, Ms=1.3):
    """The post-shock state (rho, u, v, p) that Test/PNGGeo/Global_Boudary.cpp:59-70
    holds on the left edge (its constants as defaults), in its operation order."""
    rho2 = rho1 * (gamma + 1) * Ms * Ms / (2 + (gamma - 1) * Ms * Ms)
    p2 = p1 * ((2 * gamma) * Ms * Ms / (gamma + 1) - (gamma - 1) / (gamma + 1))
    u2 = Ms * (1 - rho1 / rho2)
    return [rho2, u2, 0.0, p2]


def _num(v):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def setup(case):
    """Psy_coord + Level_Set (+ hook) + Scheme_Index + Initialize, on the CPU."""
    L = lib()
    g = case
    cells = g.M * g.N
    XY = np.zeros(cells * 6)
    U_OLD = np.zeros(cells * 4)
    U_NEW = np.zeros(cells * 4)
    bc = BOUNDS[g.Boundary]
    S = np.zeros(cells * L.or_sch_width(bc))
    L.or_psy_coord(C.c_double(g.lo_x), C.c_double(g.lo_y), C.c_double(g.L), C.c_double(g.H),
                   g.N_x, g.N_y, g.num_ghost_cell, ptr(XY))
    ls_kind = LS.get(g.case, 0)
    phi = None
    if g.phi is not None:
        ls_kind = 3
        phi = np.ascontiguousarray(g.phi, dtype=np.float64)
    prm = np.array(g.ls_prm(), dtype=np.float64)
    L.or_level_set_hook(ls_kind, ptr(prm), g.N_x, g.N_y, g.num_ghost_cell, ptr(XY), ptr(phi))
    L.or_classify(g.N_x, g.N_y, g.num_ghost_cell, ptr(XY))
    if not g.quirks:
        L.or_unit_normals(g.N_x, g.N_y, g.num_ghost_cell, ptr(XY))
    cut = None
    if g.Boundary == "CutCell":
        cut = np.zeros(cells * 8)
        L.or_cut_geometry(g.N_x, g.N_y, g.num_ghost_cell, C.c_double(g.L), C.c_double(g.H), ptr(XY), ptr(cut))
    L.or_scheme_index(bc, g.N_x, g.N_y, g.num_ghost_cell, ptr(XY), ptr(S))
    ip = np.array(g.init_prm(), dtype=np.float64)
    L.or_initialize(INIT[g.case], ptr(ip), g.N_x, g.N_y, g.num_ghost_cell, C.c_double(g.gamma),
                    ptr(U_OLD), ptr(U_NEW), ptr(XY))
    return dict(XY=XY, U_OLD=U_OLD, U_NEW=U_NEW, S=S, cut=cut)


def mask_phi(case, rgba):
    """Test/PNGGeo/Level_Set.cpp:31-149 on the CPU: phi per cell (reference
    cell order) from an (H, W, 4) uint8 image, top row first."""
    L = lib()
    g = case
    XY = np.zeros(g.M * g.N * 6)
    L.or_psy_coord(C.c_double(g.lo_x), C.c_double(g.lo_y), C.c_double(g.L), C.c_double(g.H),
                   g.N_x, g.N_y, g.num_ghost_cell, ptr(XY))
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    phi = np.zeros(g.M * g.N)
    L.or_mask_phi(img.ctypes.data_as(C.c_void_p), img.shape[1], img.shape[0], g.N_x, g.N_y, g.num_ghost_cell,
                  C.c_double(g.lo_x), C.c_double(g.lo_y), C.c_double(g.hi_x), C.c_double(g.hi_y), ptr(XY), ptr(phi))
    return phi


def mask_phi_cells(case, rgba, cells):
    """mask_phi at the listed cells only (reference cell order j*M+i); the XY
    of the grid comes from Psy_coord."""
    L = lib()
    g = case
    XY = np.zeros(g.M * g.N * 6)
    L.or_psy_coord(C.c_double(g.lo_x), C.c_double(g.lo_y), C.c_double(g.L), C.c_double(g.H),
                   g.N_x, g.N_y, g.num_ghost_cell, ptr(XY))
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    phi = np.zeros(len(cells))
    L.or_mask_phi_cells(img.ctypes.data_as(C.c_void_p), img.shape[1], img.shape[0], g.N_x, g.N_y,
                        g.num_ghost_cell, C.c_double(g.lo_x), C.c_double(g.lo_y), C.c_double(g.hi_x),
                        C.c_double(g.hi_y), ptr(XY), cells.ctypes.data_as(C.c_void_p), C.c_longlong(len(cells)),
                        ptr(phi))
    return phi


def run_oracle(case, nsteps, root_mode=0, state=None, clip=False, pvrs_stray=True):
    """Whole run like the reference driver, at most nsteps steps. Returns a dict.
    clip: the driver shortens the last step to end at Psy_time (see or_run).
    pvrs_stray: reference mode keeps the PVRS y solver's stray store
    (PVRS/Riemann_solver.cpp:550); False drops it, as the GPU path does."""
    L = lib()
    L.or_set_root_mode(int(root_mode))
    L.or_set_pvrs_stray(int(bool(pvrs_stray)))
    g = case
    st = setup(g) if state is None else state
    T = np.zeros(g.M * g.N * L.or_tmp_width(SCHEMES[g.Scheme]))
    dt_log = np.zeros(max(nsteps, 1))
    t_log = np.zeros(max(nsteps, 1))
    m_log = np.zeros(max(nsteps, 1))
    ep = np.array(g.edge_prm(), dtype=np.float64)
    L.or_set_cut(ptr(st.get("cut")))
    n = L.or_run(TIMES[g.Time], SCHEMES[g.Scheme], RIEMANNS[g.Riemann], LIMITERS[g.Limiter], BOUNDS[g.Boundary],
                 EDGES[g.edge], ptr(ep), g.quirks, C.c_double(g.L), C.c_double(g.H), g.N_x, g.N_y,
                 g.num_ghost_cell, C.c_double(g.gamma), C.c_double(g.CFL_number), C.c_double(g.Psy_time),
                 min(g.max_iter, nsteps), g.init_boundary(), ptr(st["U_OLD"]), ptr(st["U_NEW"]), ptr(st["XY"]),
                 ptr(st["S"]), ptr(T), ptr(dt_log), ptr(t_log), ptr(m_log), int(bool(clip)))
    L.or_set_root_mode(0)
    L.or_set_pvrs_stray(1)
    L.or_set_cut(None)
    st.update(steps=n, dt=dt_log[:n], t=t_log[:n], mass=m_log[:n])
    return st


def compute_dt(case, U, XY, root_mode=0):
    L = lib()
    L.or_set_root_mode(int(root_mode))
    v = L.or_compute_dt(C.c_double(case.L), C.c_double(case.H), case.N_x, case.N_y, case.num_ghost_cell,
                        C.c_double(case.gamma), C.c_double(case.CFL_number), ptr(U), ptr(XY))
    L.or_set_root_mode(0)
    return v


RECORD_W = 28  # doubles per cell of the scheme's scratch record (7 slots x 4)


def riemann_record(riemann, ydir, quirks, nx, ny, ng, gamma, rec, root_mode=0, pvrs_stray=True):
    """Riemann_solverX (ydir 0) / Riemann_solverY (1) on a whole scratch record
    (M*N*28 doubles, modified in place)."""
    L = lib()
    assert rec.size == (nx + 2 * ng) * (ny + 2 * ng) * RECORD_W
    L.or_set_root_mode(int(root_mode))
    L.or_set_pvrs_stray(int(bool(pvrs_stray)))
    L.or_riemann_record(RIEMANNS[riemann], int(ydir), int(quirks), nx, ny, ng, C.c_double(gamma), ptr(rec))
    L.or_set_root_mode(0)
    L.or_set_pvrs_stray(1)


def flux_record(gamma, rec, u_slot, f_slot=None, g_slot=None):
    """Conserve2Flux{,X,Y} on every cell of a scratch record: the state in slot
    u_slot, its x flux to slot f_slot and its y flux to g_slot (in place)."""
    n = rec.size // RECORD_W
    base = rec.ctypes.data

    def at(slot):
        return None if slot is None else C.cast(base + 8 * 4 * slot, _D)
    lib().or_flux_record(C.c_double(gamma), at(u_slot), at(f_slot), at(g_slot), C.c_longlong(n), RECORD_W)


# ------------------------------------------------------------ reference libs
def ref_lib_name(case):
    return "libref_%s_%s_%s_%s_%s.so" % (case.case, case.Scheme, case.Riemann, case.Limiter, case.Boundary)


def ref_available(case=None):
    if case is None:
        return os.path.isdir(REF_DIR) and any(f.endswith(".so") for f in os.listdir(REF_DIR))
    return os.path.exists(os.path.join(REF_DIR, ref_lib_name(case)))


_ref_cache = {}


def ref_lib(name):
    if name not in _ref_cache:
        L = C.CDLL(os.path.join(REF_DIR, name), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L.ref_run.restype = C.c_int
        L.ref_lim_L.restype = C.c_double
        L.ref_lim_L.argtypes = [C.c_double] * 3
        L.ref_lim_R.restype = C.c_double
        L.ref_lim_R.argtypes = [C.c_double] * 3
        _ref_cache[name] = L
    return _ref_cache[name]


def write_inp(case, path=None):
    if path is None:
        fd, path = tempfile.mkstemp(suffix=".inp")
        os.close(fd)
    with open(path, "w") as f:
        f.write(case.inp_text())
    return path


def ref_write_vtk(case, U, XY, iter_, now_t, directory):
    """The reference's WriteData (oracle/_ref/libref_writedata.so) on host arrays,
    writing dat_0_<iter>.vtk and appending to result.visit in `directory`."""
    L = ref_lib_writedata()
    U = np.ascontiguousarray(U, dtype=np.float64)
    XY = np.ascontiguousarray(XY, dtype=np.float64)
    here = os.getcwd()
    os.chdir(directory)
    try:
        L.ref_write_data(C.c_double(case.lo_x), C.c_double(case.lo_y), C.c_double(case.L), C.c_double(case.H),
                         case.N_x, case.N_y, case.num_ghost_cell, int(iter_), C.c_double(now_t),
                         C.c_double(case.gamma), ptr(U), ptr(XY))
    finally:
        os.chdir(here)


def ref_lib_writedata():
    name = "libref_writedata.so"
    if name not in _ref_cache:
        L = C.CDLL(os.path.join(REF_DIR, name), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L.ref_write_data.restype = None
        _ref_cache[name] = L
    return _ref_cache[name]


def ref_lib_pngmask():
    """The reference's PNG reader and mask level set (oracle/_ref/libref_pngmask.so)."""
    name = "libref_pngmask.so"
    if name not in _ref_cache:
        L = C.CDLL(os.path.join(REF_DIR, name), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L.ref_png_read.restype = C.c_int
        L.ref_mask_level_set.restype = C.c_int
        _ref_cache[name] = L
    return _ref_cache[name]


def ref_pngmask_available():
    return os.path.exists(os.path.join(REF_DIR, "libref_pngmask.so"))


def ref_png_read(path):
    """Src/Utils/PngProcess.cpp:4-72 on a file: (H, W, 4) uint8 RGBA, W, H."""
    L = ref_lib_pngmask()
    W, H = C.c_int(), C.c_int()
    cap = 1 << 26
    buf = np.zeros(cap, dtype=np.uint8)
    rc = L.ref_png_read(os.fsencode(path), buf.ctypes.data_as(C.c_void_p), C.c_longlong(cap), C.byref(W),
                        C.byref(H))
    assert rc == 0, rc
    return buf[:W.value * H.value * 4].reshape(H.value, W.value, 4).copy(), W.value, H.value


def ref_mask_level_set(directory, nx, ny, ng, extents):
    """Psy_coord + Test/PNGGeo/Level_Set.cpp:11-364 over directory/Demo.png:
    the reference's XYCOORD (M*N*6 doubles). extents = (lo_x, lo_y, hi_x, hi_y)."""
    L = ref_lib_pngmask()
    lo_x, lo_y, hi_x, hi_y = extents
    inp = os.path.join(directory, "mask.inp")
    with open(inp, "w") as f:
        f.write("lo_x = %r\nlo_y = %r\nhi_x = %r\nhi_y = %r\nN_x = %d\nN_y = %d\nnum_ghost_cell = %d\n"
                % (lo_x, lo_y, hi_x, hi_y, nx, ny, ng))
    XY = np.zeros((nx + 2 * ng) * (ny + 2 * ng) * 6)
    rc = L.ref_mask_level_set(os.fsencode(directory), os.fsencode(inp), nx, ny, ng, C.c_double(lo_x),
                              C.c_double(lo_y), C.c_double(hi_x), C.c_double(hi_y), ptr(XY))
    assert rc == 0, rc
    return XY


def run_ref(case, nsteps):
    """Run the reference translation units on the same case (edge policy and
    hooks are the reference's own)."""
    L = ref_lib(ref_lib_name(case))
    path = write_inp(case)
    try:
        cells = case.M * case.N
        out = dict(U_OLD=np.zeros(cells * 4), U_NEW=np.zeros(cells * 4), XY=np.zeros(cells * 6),
                   S=np.zeros(cells * L.ref_sch_width()))
        dt0 = C.c_double(0)
        dt_log, t_log, m_log = np.zeros(max(nsteps, 1)), np.zeros(max(nsteps, 1)), np.zeros(max(nsteps, 1))
        n = L.ref_run(path.encode(), nsteps, case.init_boundary(), ptr(out["U_OLD"]), ptr(out["U_NEW"]),
                      ptr(out["XY"]), ptr(out["S"]), C.byref(dt0), ptr(dt_log), ptr(t_log), ptr(m_log))
        out.update(steps=n, dt0=dt0.value, dt=dt_log[:n], t=t_log[:n], mass=m_log[:n])
        return out
    finally:
        os.unlink(path)


def parse_ref(path, strings=False):
    """The reference ParamReader's numbers()/strings() map for a file."""
    names = [f for f in os.listdir(REF_DIR) if f.endswith(".so")]
    L = ref_lib(sorted(names)[0])
    buf = C.create_string_buffer(1 << 16)
    n = L.ref_parse(path.encode(), int(strings), buf, len(buf))
    assert n >= 0
    out = {}
    for rec in buf.value.decode("utf-8", "surrogateescape").split("\x1e"):
        if not rec:
            continue
        k, v = rec.split("\x1f", 1)
        out[k] = v if strings else float(v)
    return out
