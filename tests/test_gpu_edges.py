"""Grid shapes at the edges of what the step kernels tile: the smallest legal
grid (4 x 4 interior cells), a strip plus one column, exactly one strip, tall
and narrow, wide and flat, with marching and row-parallel strips of several
heights: GPU == oracle bit for bit (state, every step's dt and time)."""
import pytest

from parity_util import gpu_config, gpu_hooks, nan_equal_bits

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4), (5, 4), (60, 4), (61, 5), (4, 61), (121, 7), (250, 4)]


def _run(case, nsteps, rows, par):
    import ibtfo_amd as ib
    s = ib.Solver(gpu_config(case))
    try:
        ls, init = gpu_hooks(case)
        s.geometry(ls[0], ls[1])
        s.initialize(init[0], init[1])
        s.prepare(case.init_boundary())
        s.set_tile(rows, par)
        s.step(nsteps)
        it, t, dt, rc = s.status()
        d, tl, _ = s.log(it)
        return dict(U=s.state(), steps=it, dt=d, t=tl)
    finally:
        s.close()


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_extreme_shapes_equal_oracle(oracle_mod, nx, ny):
    O = oracle_mod
    c = O.Case(N_x=nx, N_y=ny, hi_x=1.0, hi_y=ny / nx, Scheme="MUSCL", Riemann="HLL", Limiter="minmod",
               case="sodx", Psy_time=10.0)
    n = 12
    o = O.run_oracle(c, n, root_mode=1)
    assert o["steps"] == n
    print("shape %dx%d" % (nx, ny), flush=True)
    for rows, par in ((1, 0), (3, 0), (ny, 0), (1, 1), (min(ny, 7), 1)):
        g = _run(c, n, rows, par)
        assert g["steps"] == n, (rows, par)
        assert nan_equal_bits(g["U"], o["U_OLD"]) == 0, (nx, ny, rows, par)
        assert nan_equal_bits(g["dt"], o["dt"]) == 0 and nan_equal_bits(g["t"], o["t"]) == 0, (rows, par)
