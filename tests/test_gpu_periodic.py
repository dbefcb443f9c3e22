"""Periodic edges with no body: the step kernel writes the ring itself (each
interior cell within ng of an edge also stores its images across the x edge,
the y edge and the corner), and an RK2 step runs its two stage kernels back to
back (stage 2 reads U1 from the buffer stage 1 wrote, and the step's finish
keeps the current buffer). GPU == oracle bit for bit, the whole state with its
ring and the dt log, for both kernel forms (marching strips and row-parallel),
single step and RK2, eager and graph-replayed, on grids where every column
near the x edges has two images (N_x < 2 ng: ng = 3 with N_x = 4, 5; N_x = ng =
4), and with ragged last strips (ng = 2). The reference ring on the Regular
boundary (also written by the step kernel) runs its RK2 stages back to back
too: the same checks on an x-split Sod tube (MUSCL + minmod in reference mode,
MHM + van Leer in corrected mode)."""
import pytest

from parity_util import gpu_config, gpu_hooks, nan_equal_bits

pytestmark = pytest.mark.gpu

SHAPES = [(4, 6, 3), (5, 8, 3), (4, 8, 4), (61, 37, 2), (130, 27, 2)]


def _case(O, nx, ny, ng, time):
    return O.Case(N_x=nx, N_y=ny, num_ghost_cell=ng, lo_x=0.0, lo_y=0.0, hi_x=10.0, hi_y=10.0, Psy_time=1e30,
                  max_iter=10 ** 9, CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Time=time, case="shu",
                  edge="periodic", quirks=0, extra=dict(x0=5.0, y0=5.0, beta=5.0))


def _run(case, nsteps, rows, par, graph_chunk=0):
    import ibtfo_amd as ib
    s = ib.Solver(gpu_config(case))
    try:
        ls, init = gpu_hooks(case)
        s.geometry(ls[0], ls[1])
        s.initialize(init[0], init[1])
        s.prepare(case.init_boundary())
        s.set_tile(rows, par)
        s.step(nsteps, graph_chunk)
        it, t, dt, rc = s.status()
        d, tl, m = s.log(min(it, 4000))
        return dict(U=s.state(), steps=it, dt_log=d)
    finally:
        s.close()


@pytest.mark.parametrize("time", ["SingleStep", "RK2"])
@pytest.mark.parametrize("nx,ny,ng", SHAPES)
def test_periodic_ring_in_step_kernel(oracle_mod, nx, ny, ng, time):
    O = oracle_mod
    c = _case(O, nx, ny, ng, time)
    n = 12
    o = O.run_oracle(c, n, root_mode=1)
    assert o["steps"] == n
    for rows, par, chunk in ((4, 0, 0), (3, 1, 0), (7, 0, 5), (2, 1, 5)):
        print("periodic ring:", nx, ny, ng, time, "rows", rows, "row-parallel", par, "graph chunk", chunk, flush=True)
        g = _run(c, n, rows, par, chunk)
        assert g["steps"] == n, (rows, par, chunk)
        assert nan_equal_bits(g["U"], o["U_OLD"]) == 0, (rows, par, chunk)
        assert nan_equal_bits(g["dt_log"], o["dt"]) == 0, (rows, par, chunk)


@pytest.mark.parametrize("limiter,scheme", [("minmod", "MUSCL"), ("VanLeer", "MHM")])
def test_rk2_reference_ring_in_place(oracle_mod, limiter, scheme):
    O = oracle_mod
    c = O.Case(N_x=130, N_y=27, hi_y=0.25, Scheme=scheme, Limiter=limiter, Time="RK2", case="sodx",
               quirks=1 if scheme == "MUSCL" else 0)
    n = 20
    o = O.run_oracle(c, n, root_mode=1)
    assert o["steps"] == n
    for rows, par, chunk in ((4, 0, 0), (3, 1, 0), (7, 0, 6), (2, 1, 6)):
        print("reference ring RK2:", scheme, limiter, "rows", rows, "row-parallel", par, "graph chunk", chunk,
              flush=True)
        g = _run(c, n, rows, par, chunk)
        assert g["steps"] == n, (rows, par, chunk)
        assert nan_equal_bits(g["U"], o["U_OLD"]) == 0, (rows, par, chunk)
        assert nan_equal_bits(g["dt_log"], o["dt"]) == 0, (rows, par, chunk)
