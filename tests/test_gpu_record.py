"""The scheme's inner operators on the reference's scratch record, on the GPU:
ibtfo_riemann_record(_host) (Riemann_solverX/Y, Riemann_solver.H:22-42) and
ibtfo_flux_record(_host) (Conserve2Flux{,X,Y}, Conserve2Flux.H:11-37), and the
same through the reference signatures (tests/drivers/ref_record_driver), against
the oracle on whole seeded records (every slot compared, NaNs equal):

  * HLL, HLLC and Steger-Warming, both directions, reference and corrected mode:
    bit for bit against the oracle with sqrt sound speeds;
  * PVRS: the slots it does not write bit for bit, its face fluxes within the
    FP64 tolerance (its pow() calls use the device libm), the stray store of
    PVRS/Riemann_solver.cpp:550 dropped as in the step kernel;
  * Conserve2Flux{,X,Y}: bit for bit, in place (state and flux in one slot) too;
  * the device-pointer entry points give the host entry points' bits.

The oracle's record operators equal the reference's own translation units bit
for bit (tests/test_oracle_record.py)."""
import os
import subprocess

import numpy as np
import pytest

from parity_util import REL_TOL, nan_equal_bits, rel_linf
from record_util import RECORD_W, SLOT_HI, SLOT_HI_FLUX, SLOT_LO, SLOT_LO_FLUX, SLOT_OLD, random_record

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "drivers", "ref_record_driver")
NX, NY, NG, GAMMA = 61, 37, 2, 1.4


class _DeviceCopy:
    """A device copy of a host array through the HIP runtime the library is
    linked against (libamdhip64.so.7); no second runtime enters the process."""

    def __init__(self, host):
        import ctypes as C
        import ibtfo_amd as ib
        ib.lib()
        self.C, self.hip = C, C.CDLL("libamdhip64.so.7")
        self.n = host.nbytes
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(self.n)) == 0
        assert self.hip.hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.n), 1) == 0  # H2D

    def get(self):
        out = np.empty(self.n // 8)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.p, self.C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _solver(riemann, quirks):
    import ibtfo_amd as ib
    return ib.Solver(ib.config(N_x=NX, N_y=NY, num_ghost_cell=NG, gamma=GAMMA, Riemann=riemann, quirks=quirks))


def _expect(O, riemann, ydir, quirks, rec):
    o = rec.copy()
    O.riemann_record(riemann, ydir, quirks, NX, NY, NG, GAMMA, o, root_mode=1, pvrs_stray=False)
    return o


def _compare(riemann, g, o):
    if riemann != "PVRS":
        assert nan_equal_bits(g, o) == 0
        return
    gv, ov = g.reshape(-1, 7, 4), o.reshape(-1, 7, 4)
    rest = [s for s in range(7) if s != SLOT_OLD]
    assert nan_equal_bits(gv[:, rest], ov[:, rest]) == 0
    assert (np.isnan(gv[:, SLOT_OLD]) == np.isnan(ov[:, SLOT_OLD])).all()
    assert max(rel_linf(gv[:, SLOT_OLD], ov[:, SLOT_OLD])) <= REL_TOL


@pytest.mark.parametrize("quirks", [1, 0])
@pytest.mark.parametrize("ydir", [0, 1])
@pytest.mark.parametrize("riemann", ["HLL", "HLLC", "PVRS", "StegerWarming"])
def test_riemann_record_equals_oracle(oracle_mod, riemann, ydir, quirks):
    O = oracle_mod
    print("riemann record:", riemann, ydir, quirks, flush=True)
    rec = random_record(NX, NY, NG, seed=200 + 2 * ydir + quirks)
    o = _expect(O, riemann, ydir, quirks, rec)
    s = _solver(riemann, quirks)
    try:
        g = rec.copy()
        s.riemann_record(g, ydir)
    finally:
        s.close()
    _compare(riemann, g, o)
    assert nan_equal_bits(g, rec) > 0


@pytest.mark.parametrize("riemann", ["HLL", "PVRS"])
def test_riemann_record_device_equals_host(riemann):
    rec = random_record(NX, NY, NG, seed=31)
    s = _solver(riemann, 1)
    d = None
    try:
        h = rec.copy()
        s.riemann_record(h, 1)
        d = _DeviceCopy(rec)
        s.riemann_record_device(d.p.value, 1)
        g = d.get()
    finally:
        if d is not None:
            d.free()
        s.close()
    assert nan_equal_bits(g, h) == 0


@pytest.mark.parametrize("u_slot,f_slot,g_slot", [(SLOT_LO, SLOT_LO_FLUX, None), (SLOT_HI, None, SLOT_HI_FLUX),
                                                  (SLOT_LO, SLOT_LO_FLUX, 0), (SLOT_LO, SLOT_LO, None)])
def test_flux_record_equals_oracle(oracle_mod, u_slot, f_slot, g_slot):
    import ibtfo_amd as ib
    O = oracle_mod
    rec = random_record(NX, NY, NG, seed=41)
    o = rec.copy()
    O.flux_record(GAMMA, o, u_slot, f_slot, g_slot)
    h = rec.copy()
    ib.flux_record(GAMMA, h, u_slot, f_slot, g_slot)
    assert nan_equal_bits(h, o) == 0
    d = _DeviceCopy(rec)
    try:
        base, n = d.p.value, rec.size // RECORD_W

        def at(slot):
            return None if slot is None else base + 32 * slot
        ib.flux_record_device(GAMMA, at(u_slot), at(f_slot), at(g_slot), n, RECORD_W)
        g = d.get()
    finally:
        d.free()
    assert nan_equal_bits(g, o) == 0


def test_flux_record_packed_states(oracle_mod):
    """Stride 4 (states packed one after another) and a zero-length call."""
    import ctypes as C
    import ibtfo_amd as ib
    O = oracle_mod
    rec = random_record(NX, NY, NG, seed=61).reshape(-1, 7, 4)
    u = np.ascontiguousarray(rec[:, SLOT_HI])
    n = u.shape[0]
    f_o, g_o = np.zeros_like(u), np.zeros_like(u)
    O.lib().or_flux_record(C.c_double(GAMMA), O.ptr(u), O.ptr(f_o), O.ptr(g_o), C.c_longlong(n), 4)
    f_g, g_g = np.zeros_like(u), np.zeros_like(u)
    L = ib.lib()
    assert L.ibtfo_flux_record_host(C.c_double(GAMMA), ib.ptr(u), ib.ptr(f_g), ib.ptr(g_g), C.c_longlong(n), 4) == ib.OK
    assert nan_equal_bits(f_g, f_o) == 0 and nan_equal_bits(g_g, g_o) == 0
    assert L.ibtfo_flux_record_host(C.c_double(GAMMA), ib.ptr(u), ib.ptr(f_g), None, C.c_longlong(0), 4) == ib.OK


INP = """N_x = {nx}
N_y = {ny}
num_ghost_cell = {ng}
gamma = {gamma!r}
Scheme = MUSCL
Riemann = {riemann}
Limiter = minmod
Boundary = Regular
Time = SingleStep
"""


@pytest.mark.parametrize("riemann", ["HLL", "StegerWarming"])
def test_reference_signature_record_driver(oracle_mod, tmp_path, riemann):
    """Riemann_solverX/Y and Conserve2Flux{,X,Y} called the way the reference's
    schemes call them, by a driver linked against the library's reference
    signatures, on a host U_TMP: bit for bit the oracle's (reference mode)."""
    O = oracle_mod
    (tmp_path / "case.inp").write_text(INP.format(nx=NX, ny=NY, ng=NG, gamma=GAMMA, riemann=riemann))
    rec = random_record(NX, NY, NG, seed=51)
    rec.tofile(tmp_path / "rec.bin")
    out = tmp_path / "out"
    out.mkdir()
    print("record driver:", riemann, flush=True)
    r = subprocess.run([DRIVER, "case.inp", "rec.bin", str(out)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = {k: np.fromfile(out / (k + ".bin")) for k in ["x", "y", "fx", "fy", "f"]}
    for ydir, k in ((0, "x"), (1, "y")):
        assert nan_equal_bits(got[k], _expect(O, riemann, ydir, 1, rec)) == 0, k
    for k, args in (("fx", (SLOT_LO, SLOT_LO_FLUX, None)), ("fy", (SLOT_HI, None, SLOT_HI_FLUX)),
                    ("f", (SLOT_LO, SLOT_LO_FLUX, 0))):
        o = rec.copy()
        O.flux_record(GAMMA, o, *args)
        assert nan_equal_bits(got[k], o) == 0, k


def test_reference_signature_lax_record(oracle_mod, tmp_path):
    """Conserve2Flux{,X,Y} called by a driver whose .inp selects Lax: the
    library takes Lax's 16-double record (state in slot 0, x flux to slot 2, y
    flux to slot 3, Lax_Scheme.cpp:41): bit for bit the oracle's."""
    import ctypes as C
    O = oracle_mod
    inp = INP.format(nx=NX, ny=NY, ng=NG, gamma=GAMMA, riemann="HLL").replace("Scheme = MUSCL", "Scheme = Lax")
    (tmp_path / "case.inp").write_text(inp)
    rec = random_record(NX, NY, NG, seed=71).reshape(-1, 7, 4)[:, [5, 1, 2, 3]].copy().reshape(-1)
    rec.tofile(tmp_path / "rec.bin")
    out = tmp_path / "out"
    out.mkdir()
    print("record driver: Lax", flush=True)
    r = subprocess.run([DRIVER, "case.inp", "rec.bin", str(out)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    n = rec.size // 16

    def at(arr, slot):
        return C.cast(arr.ctypes.data + 32 * slot, C.POINTER(C.c_double))
    for k, (fs, gs) in (("fx", (2, None)), ("fy", (None, 3)), ("f", (2, 3))):
        o = rec.copy()
        O.lib().or_flux_record(C.c_double(GAMMA), at(o, 0), at(o, fs) if fs else None, at(o, gs) if gs else None,
                               C.c_longlong(n), 16)
        got = np.fromfile(out / (k + ".bin"))
        assert got.size == rec.size and nan_equal_bits(got, o) == 0, k
    assert not (out / "x.bin").exists()


def test_riemann_record_arguments(oracle_mod):
    import ibtfo_amd as ib
    s = _solver("HLL", 1)
    try:
        rec = random_record(NX, NY, NG, seed=1)
        for d in (-1, 2):
            assert ib.lib().ibtfo_riemann_record_host(s.h, d, ib.ptr(rec)) == ib.E_ARG
        assert ib.lib().ibtfo_riemann_record_host(s.h, 0, None) == ib.E_ARG
    finally:
        s.close()
