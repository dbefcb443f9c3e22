"""Oracle against the reference's translation units built from source in this
container (oracle/_ref); skipped where they are not built. Larger and longer
runs than the committed fixtures, same bit-exact rule (NaNs equal)."""
import numpy as np
import pytest

from parity_util import nan_equal_bits


def _cmp(O, c, n):
    if not O.ref_available(c):
        pytest.skip("oracle/_ref not built")
    a = O.run_oracle(c, n)
    b = O.run_ref(c, n)
    assert a["steps"] == b["steps"]
    for k in ["U_OLD", "XY", "S"]:
        assert nan_equal_bits(a[k], b[k]) == 0, k
    assert nan_equal_bits(a["dt"], b["dt"]) == 0
    assert nan_equal_bits(a["t"], b["t"]) == 0


# Lax uses neither a Riemann solver nor a limiter: one combination covers it
_COMBOS = [(s, r, l) for s in ["MUSCL", "MHM", "Godunov"] for r in ["HLL", "PVRS", "StegerWarming"]
           for l in ["minmod", "VanLeer"]] + [("Lax", "HLL", "minmod")]


@pytest.mark.parametrize("scheme,rs,lim", _COMBOS)
def test_sodx_combos(oracle_mod, scheme, rs, lim):
    O = oracle_mod
    c = O.Case(N_x=64, N_y=16, hi_y=0.25, Scheme=scheme, Riemann=rs, Limiter=lim, case="sodx")
    _cmp(O, c, 40)


@pytest.mark.parametrize("bound", ["IDWBaseGFM", "LinearBaseGFM", "RiemannBaseGFM", "BCGFM", "HBCGFM"])
def test_incline_gfm(oracle_mod, bound):
    O = oracle_mod
    c = O.Case(N_x=64, N_y=64, hi_y=1.0, Boundary=bound, case="incline", Psy_time=0.3)
    _cmp(O, c, 40)


@pytest.mark.parametrize("case,n", [("incline", 64), ("incline", 100), ("vortex", 48)])
def test_simplebound(oracle_mod, case, n):
    """SimpleBound (axis-direction donors, in-place sequential fill) against the
    reference's SimpleBound sources built through the signature adapter."""
    O = oracle_mod
    kw = dict(Psy_time=0.3) if case == "incline" else dict(extra=dict(fa=1.0, f0=0.0, f1=1.0))
    _cmp(O, O.Case(N_x=n, N_y=n, hi_y=1.0, Boundary="SimpleBound", case=case, **kw), 40)


@pytest.mark.parametrize("rs", ["StegerWarming", "HLL"])
def test_pnggeo_hooks(oracle_mod, rs):
    """Test/PNGGeo's Mach 1.3 shock initialiser and its edge hook (single-device
    signatures, adapted), over InclineSod's channel, IDW ghost fluid."""
    O = oracle_mod
    c = O.Case(N_x=64, N_y=64, hi_y=1.0, Boundary="IDWBaseGFM", Riemann=rs, case="pngshock", edge="pnggeo",
               Psy_time=0.3, CFL_number=0.1, extra=dict(rho_L=1.0, u_L=0.0, v_L=0.0, p_L=0.71429))
    _cmp(O, c, 40)


def test_sod_400x100_full_run(oracle_mod):
    """The 400x100 Sod tube (MHM + minmod + HLLC) with the reference's row-split Sod hook replaced by the
    x-split hook (the same hook compiled into both sides)."""
    O = oracle_mod
    c = O.Case(N_x=400, N_y=100, hi_y=0.25, Scheme="MHM", Riemann="HLLC", case="sodx", CFL_number=0.5,
               Psy_time=0.2)
    c.Riemann = "HLL"  # the reference's HLLC directory is HLL's code; built once as HLL
    _cmp(O, c, 100000)


def test_vortex_and_rows(oracle_mod):
    O = oracle_mod
    _cmp(O, O.Case(N_x=48, N_y=48, hi_y=1.0, Boundary="IDWBaseGFM", case="vortex",
                   extra=dict(fa=1.0, f0=0.0, f1=1.0)), 30)
    _cmp(O, O.Case(N_x=40, N_y=40, hi_y=1.0, case="sod"), 30)


def _collision(O, scheme, n):
    return O.Case(N_x=n, N_y=n, hi_y=1.0, Scheme=scheme, Riemann="PVRS", case="sod", CFL_number=0.3,
                  extra=dict(rho_L=1.0, u_L=0.3, v_L=0.0, p_L=1.0, rho_R=1.0, u_R=0.0, v_R=-2.0, p_R=2.0))


@pytest.mark.parametrize("scheme", ["MUSCL", "MHM"])
def test_pvrs_2d_inclined_channel(oracle_mod, scheme):
    """PVRS with its y solver on the inclined channel (IDW ghost fluid)."""
    O = oracle_mod
    c = O.Case(N_x=64, N_y=64, hi_y=1.0, Boundary="IDWBaseGFM", Scheme=scheme, Riemann="PVRS", case="incline",
               Psy_time=0.3)
    _cmp(O, c, 40)


@pytest.mark.parametrize("scheme", ["MUSCL", "MHM"])
def test_pvrs_2d_stray_store(oracle_mod, scheme):
    """A supersonic collision along y takes the y solver's branch whose energy
    flux the reference stores through the state index macro
    (PVRS/Riemann_solver.cpp:550): into the scratch record of another cell.
    The oracle reproduces it bit for bit (NaN positions included: the
    reference's PVRS fan states go non-finite in this flow). The store lands in
    flux slots and changes the result from the first step on: with it dropped
    (the GPU path's definition) the oracle differs from the reference, first in
    the energy of interior row 3 (then copied into the ring by Global_Boundary)."""
    import ctypes as C
    O = oracle_mod
    c = _collision(O, scheme, 40)
    _cmp(O, c, 8)
    L = O.lib()
    f, g = C.c_longlong(), C.c_longlong()
    L.or_pvrs_stray_counts(C.byref(f), C.byref(g))
    a = O.run_oracle(c, 1)
    L.or_pvrs_stray_counts(C.byref(f), C.byref(g))
    assert f.value > 0 and g.value > 0
    b = O.run_oracle(c, 1, pvrs_stray=False)
    d = np.flatnonzero(a["U_OLD"] != b["U_OLD"])
    assert len(d) > 0
    rows = sorted({int(q // 4 // c.M) for q in d})
    assert rows[0] in (0, 1, 2, 3) and all(q % 4 == 3 for q in d)


@pytest.mark.parametrize("case", ["sod", "incline"])
def test_lax_2d(oracle_mod, case):
    """Lax (its own 16-double scratch record, no Riemann solver) in 2-D: the
    row-split Sod hook (Regular) and the inclined channel with IDW ghost fluid."""
    O = oracle_mod
    kw = dict(Boundary="IDWBaseGFM", Psy_time=0.3) if case == "incline" else {}
    _cmp(O, O.Case(N_x=64, N_y=64, hi_y=1.0, Scheme="Lax", case=case, **kw), 40)


@pytest.mark.parametrize("scheme,rs,lim", [("Godunov", "HLL", "minmod"), ("Godunov", "StegerWarming", "minmod"),
                                           ("MUSCL", "HLL", "VanLeer"), ("MHM", "StegerWarming", "minmod"),
                                           ("MHM", "HLL", "VanLeer")])
def test_incline_idw_combos(oracle_mod, scheme, rs, lim):
    """Scheme / solver / limiter combinations in 2-D on the IDW channel,
    reference mode (the van Leer macro expansion included)."""
    O = oracle_mod
    _cmp(O, O.Case(N_x=64, N_y=64, hi_y=1.0, Boundary="IDWBaseGFM", Scheme=scheme, Riemann=rs, Limiter=lim,
                   case="incline", Psy_time=0.3), 40)
