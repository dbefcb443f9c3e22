"""The row-parallel step kernel (ibtfo_set_tile(ctx, rows, 1): each row of a
strip on its own wave, for small grids where the marching strip's rows set a
step's latency) gives the marching strip's state bit for bit, and the oracle's
where the strip is pinned to it: every scheme (MUSCL, MHM, Godunov), flux (HLL,
Steger-Warming, PVRS), limiter, boundary family (Regular with the fused
reference ring, IDW and BCGFM ghost fluid, SimpleBound, cut cells), the
reference-mode quirks, Heun's RK2 and the mixed-precision form, on grids whose
last strips are ragged in x and y; and the argument checks (1..7 rows, no
Lax)."""
import numpy as np
import pytest

from parity_util import gpu_config, gpu_hooks, nan_equal_bits

pytestmark = pytest.mark.gpu


def _solver(case, precision=None):
    import ibtfo_amd as ib
    cfg = gpu_config(case)
    if precision is not None:
        cfg.precision = ib.PRECISION[precision]
    s = ib.Solver(cfg)
    ls, init = gpu_hooks(case)
    if case.phi is not None:
        s.geometry("field", (), phi=case.phi)
    else:
        s.geometry(ls[0], ls[1])
    s.initialize(init[0], init[1])
    s.prepare(case.init_boundary())
    return s


def _run(case, nsteps, rows, par, precision=None, graph_chunk=0):
    s = _solver(case, precision)
    try:
        s.set_tile(rows, par)
        assert s.tile_form() == bool(par) and s.tile()[0] == rows
        s.step(nsteps, graph_chunk)
        it, t, dt, rc = s.status()
        d, tl, m = s.log(min(it, 4000))
        return dict(U=s.state(), steps=it, t=t, dt=dt, rc=rc, dt_log=d, t_log=tl, mass=m)
    finally:
        s.close()


def _bodies(O, nx, ny, boundary, riemann="HLL"):
    from ibtfo_amd import workloads as W
    c = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=10 ** 9,
               CFL_number=0.5, Scheme="MUSCL", Riemann=riemann, Limiter="VanLeer", Boundary=boundary,
               case="uniform", edge="inflow", quirks=0, extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    dx = 2.0 / nx
    Li = -dx * 2 + dx / 2
    x = (Li + np.arange(c.M) * dx)[None, :].repeat(c.N, 0)
    y = (Li + np.arange(c.N) * dx)[:, None].repeat(c.M, 1)
    c.phi = W.shapes_phi(W.bodies_shapes(), x, y).reshape(-1)
    return c


CASES = {
    # 130 columns: strips of 60, 60, 10; 27 rows: ragged last strips for every height
    "sod_mhm_hllc_minmod_ref": dict(N_x=130, N_y=27, hi_y=0.25, Scheme="MHM", Riemann="HLLC", Limiter="minmod",
                                    case="sodx"),
    "sod_godunov_sw_ref": dict(N_x=130, N_y=27, hi_y=0.25, Scheme="Godunov", Riemann="StegerWarming",
                               Limiter="minmod", case="sodx"),
    "sod_muscl_vanalbada_corrected": dict(N_x=130, N_y=27, hi_y=0.25, Scheme="MUSCL", Riemann="HLL",
                                          Limiter="VanAlbada", case="sodx", quirks=0),
    "incline_muscl_vanleer_idw_ref": dict(N_x=97, N_y=61, hi_y=1.0, Scheme="MUSCL", Riemann="HLL",
                                          Limiter="VanLeer", Boundary="IDWBaseGFM", case="incline", Psy_time=0.3),
    "incline_mhm_sw_bcgfm_ref": dict(N_x=97, N_y=61, hi_y=1.0, Scheme="MHM", Riemann="StegerWarming",
                                     Limiter="minmod", Boundary="BCGFM", case="incline", Psy_time=0.3),
    "incline_muscl_pvrs_idw_ref": dict(N_x=97, N_y=61, hi_y=1.0, Scheme="MUSCL", Riemann="PVRS", Limiter="minmod",
                                       Boundary="IDWBaseGFM", case="incline", Psy_time=0.3),
    "annulus_simplebound_ref": dict(N_x=64, N_y=64, hi_y=1.0, Boundary="SimpleBound", case="vortex",
                                    extra=dict(fa=1.0, f0=0.0, f1=1.0)),
    "shu_rk2_periodic": dict(N_x=70, N_y=70, lo_x=0.0, lo_y=0.0, hi_x=10.0, hi_y=10.0, Psy_time=1e30,
                             Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Time="RK2", case="shu",
                             edge="periodic", quirks=0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_rowpar_equals_strip(oracle_mod, name):
    O = oracle_mod
    c = O.Case(**CASES[name])
    n = 20
    print("row-parallel vs strip:", name, flush=True)
    for rows in (1, 4, 7):
        a = _run(c, n, rows, 0)
        b = _run(c, n, rows, 1)
        assert a["steps"] == b["steps"] == n, (name, rows)
        assert nan_equal_bits(a["U"], b["U"]) == 0, (name, rows)
        assert nan_equal_bits(a["dt_log"], b["dt_log"]) == 0 and nan_equal_bits(a["t_log"], b["t_log"]) == 0
        # the mass diagnostic: the same terms grouped differently (NaN where the state holds NaN)
        assert np.allclose(a["mass"], b["mass"], rtol=1e-13, atol=0, equal_nan=True), (name, rows)


@pytest.mark.parametrize("boundary,riemann", [("IDWBaseGFM", "HLL"), ("CutCell", "HLL"),
                                              ("CutCell", "StegerWarming")])
def test_rowpar_bodies_equal_oracle(oracle_mod, boundary, riemann):
    """Mach 3 past the seeded bodies, 250x122 (corrected mode; cut cells with
    small-cell merging): the row-parallel kernel == the oracle bit for bit."""
    O = oracle_mod
    c = _bodies(O, 250, 122, boundary, riemann)
    o = O.run_oracle(c, 15, root_mode=1)
    for rows in (2, 5):
        print("row-parallel bodies:", boundary, riemann, rows, flush=True)
        g = _run(c, 15, rows, 1)
        assert g["steps"] == o["steps"] == 15
        assert nan_equal_bits(g["U"], o["U_OLD"]) == 0, rows
        assert nan_equal_bits(g["dt_log"], o["dt"]) == 0


def test_rowpar_mixed_and_graph(oracle_mod):
    """The mixed-precision form and graph replay: row-parallel == strip."""
    O = oracle_mod
    c = _bodies(O, 250, 122, "IDWBaseGFM")
    a = _run(c, 30, 3, 0, precision="mixed", graph_chunk=10)
    b = _run(c, 30, 3, 1, precision="mixed", graph_chunk=10)
    assert a["steps"] == b["steps"] == 30
    assert nan_equal_bits(a["U"], b["U"]) == 0


def test_set_tile_arguments(oracle_mod):
    import ibtfo_amd as ib
    O = oracle_mod
    s = _solver(O.Case(**CASES["sod_mhm_hllc_minmod_ref"]))
    try:
        for rows, par in ((0, 1), (8, 1), (-1, 0), (3, 2)):
            assert ib.lib().ibtfo_set_tile(s.h, rows, par) == ib.E_ARG, (rows, par)
        s.set_tile(7, 1)
        assert s.tile_form() and s.tile()[0] == 7
        s.set_rows(5)  # back to the marching strip
        assert not s.tile_form() and s.tile()[0] == 5
    finally:
        s.close()
    lax = O.Case(N_x=64, N_y=16, hi_y=0.25, Scheme="Lax", case="sodx")
    s = _solver(lax)
    try:
        with pytest.raises(ib.IbtfoError):
            s.set_tile(2, 1)
    finally:
        s.close()
