"""The oracle's record operators against the reference's own Riemann_solverX/Y
(HLL, HLLC, PVRS and Steger-Warming translation units) and Conserve2Flux{,X,Y},
compiled from source in this container (oracle/_ref): bit for bit on whole
seeded scratch records, every slot compared (NaNs equal), PVRS's stray store
included. Skipped where oracle/_ref is not built."""
import ctypes as C

import numpy as np
import pytest

from parity_util import nan_equal_bits
from record_util import RECORD_W, SLOT_HI, SLOT_HI_FLUX, SLOT_LO, SLOT_LO_FLUX, random_record

NX, NY, NG, GAMMA = 61, 37, 2, 1.4


def _ref(O, riemann):
    c = O.Case(N_x=64, N_y=16, hi_y=0.25, Scheme="MUSCL", Riemann=riemann, Limiter="minmod", case="sodx")
    if not O.ref_available(c):
        pytest.skip("oracle/_ref not built")
    return O.ref_lib(O.ref_lib_name(c))


@pytest.mark.parametrize("riemann", ["HLL", "HLLC", "PVRS", "StegerWarming"])
@pytest.mark.parametrize("ydir", [0, 1])
def test_riemann_record_equals_reference(oracle_mod, riemann, ydir):
    O = oracle_mod
    R = _ref(O, riemann)
    rec = random_record(NX, NY, NG, seed=100 + ydir)
    a = rec.copy()
    b = rec.copy()
    O.lib().or_pvrs_stray_counts(C.byref(C.c_longlong()), C.byref(C.c_longlong()))
    O.riemann_record(riemann, ydir, 1, NX, NY, NG, GAMMA, a)
    faces, gold = C.c_longlong(), C.c_longlong()
    O.lib().or_pvrs_stray_counts(C.byref(faces), C.byref(gold))
    U_OLD = np.zeros((NX + 2 * NG) * (NY + 2 * NG) * 4)
    R.ref_riemann_record(ydir, NX, NY, NG, C.c_double(GAMMA), O.ptr(U_OLD), O.ptr(b))
    assert nan_equal_bits(a, b) == 0
    written = (a.view(np.uint64) != rec.view(np.uint64)).reshape(-1, RECORD_W)[:, 8:12].any(axis=1)
    if riemann in ("HLL", "HLLC") and ydir:
        # the y solver stores nothing when S_D >= 0 (HLL/Riemann_solver.cpp:224-226)
        assert 0.05 < 1 - written.mean() < 0.5
    else:
        assert written.mean() > 0.98  # the face flux slot is written
    if riemann == "PVRS" and ydir:
        assert faces.value > 0  # the stray store's branch was taken (and matched)
    if riemann in ("HLL", "HLLC"):
        # the branches the faces took: upwind either way and the HLL average
        f = a.reshape(-1, 7, 4)
        fin = np.isfinite(f[:, 2]).all(axis=1)
        up_lo = (f[:, 2] == f[:, SLOT_LO_FLUX]).all(axis=1) & fin
        up_hi = (f[:, 2] == f[:, SLOT_HI_FLUX]).all(axis=1) & fin
        if not ydir:
            assert up_lo.sum() > 50 and up_hi.sum() > 50 and (~up_lo & ~up_hi & fin).sum() > 50
    assert np.isnan(a).any()  # the non-physical states reached the comparison


@pytest.mark.parametrize("which", [0, 1, 2])
def test_conserve2flux_record_equals_reference(oracle_mod, which):
    O = oracle_mod
    R = _ref(O, "HLL")
    rec = random_record(NX, NY, NG, seed=7)
    a = rec.copy()
    b = rec.copy()
    u, f, g = SLOT_LO, SLOT_LO_FLUX, SLOT_HI_FLUX
    O.flux_record(GAMMA, a, u, f if which in (0, 2) else None, g if which in (1, 2) else None)

    def at(slot):
        return C.cast(b.ctypes.data + 8 * 4 * slot, C.POINTER(C.c_double))
    R.ref_conserve2flux(which, NX, NY, NG, C.c_double(GAMMA), at(u), at(f), at(g))
    assert nan_equal_bits(a, b) == 0
    assert (a != rec).sum() > 0


def test_flux_record_closed_form(oracle_mod):
    """Conserve2Flux against the textbook Euler flux, evaluated independently in
    numpy: F = (rho u, rho u^2 + p, rho u v, (E + p) u), G with u and v swapped
    roles (relative 1e-15 on physical states)."""
    O = oracle_mod
    rec = random_record(20, 10, 2, seed=3, bad=False)
    O.flux_record(GAMMA, rec, SLOT_HI, SLOT_LO_FLUX, SLOT_HI_FLUX)
    r = rec.reshape(-1, 7, 4)
    U = r[:, SLOT_HI]
    rho, u, v, E = U[:, 0], U[:, 1] / U[:, 0], U[:, 2] / U[:, 0], U[:, 3]
    p = (GAMMA - 1) * (E - 0.5 * rho * (u * u + v * v))
    F = np.stack([rho * u, rho * u * u + p, rho * u * v, (E + p) * u], axis=1)
    G = np.stack([rho * v, rho * u * v, rho * v * v + p, (E + p) * v], axis=1)
    scale = np.abs(F).max(axis=0)
    assert (np.abs(r[:, SLOT_LO_FLUX] - F).max(axis=0) <= 1e-15 * 8 * scale).all()
    scale = np.abs(G).max(axis=0)
    assert (np.abs(r[:, SLOT_HI_FLUX] - G).max(axis=0) <= 1e-15 * 8 * scale).all()


@pytest.mark.parametrize("which", [0, 1, 2])
def test_conserve2flux_lax_record_equals_reference(oracle_mod, which):
    """Lax's 16-double record (Lax/EQDefine.H:5-6): the state in slot 0, its x
    flux to slot 2 and its y flux to slot 3, as Lax_Scheme.cpp:41 calls it,
    against the reference's Conserve2Flux built with that record width."""
    O = oracle_mod
    c = O.Case(N_x=64, N_y=16, hi_y=0.25, Scheme="Lax", case="sodx")
    if not O.ref_available(c):
        pytest.skip("oracle/_ref not built")
    R = O.ref_lib(O.ref_lib_name(c))
    assert R.ref_tmp_width() == 16
    rec = random_record(NX, NY, NG, seed=9).reshape(-1, 7, 4)[:, [5, 1, 2, 3]].copy().reshape(-1)
    a, b = rec.copy(), rec.copy()
    n = rec.size // 16

    def at(arr, slot):
        return C.cast(arr.ctypes.data + 8 * 4 * slot, C.POINTER(C.c_double))
    f = which in (0, 2)
    g = which in (1, 2)
    O.lib().or_flux_record(C.c_double(GAMMA), at(a, 0), at(a, 2) if f else None, at(a, 3) if g else None,
                           C.c_longlong(n), 16)
    R.ref_conserve2flux(which, NX, NY, NG, C.c_double(GAMMA), at(b, 0), at(b, 2), at(b, 3))
    assert nan_equal_bits(a, b) == 0
    assert (a != rec).sum() > 0
