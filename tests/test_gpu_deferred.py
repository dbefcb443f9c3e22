"""The deferred finish of small grids. With row-parallel strips, a single-step
solve on the Regular boundary with the reference edge hook (or periodic edges,
whose ring the step kernel writes too) and at most 256 blocks runs each step as
one kernel: every block of kernel k reduces kernel
k-1's partial records and advances the clock itself, and a flush kernel closes
each batch of steps (an ibtfo_step call, a graph replay). The run must equal the
two-kernel form (a finish kernel after each step) bit for bit: state, every
step's dt and time, the step count and the status; the mass log to the
rounding of its per-block grouping (marching and row-parallel strips sum a
block's cells in different orders). Eager and graph-replayed batches (even and
odd chunks, eager tails), a run ending inside a batch and stepped past its end,
the caller's dt (ibtfo_step_dt) and per-launch timing are covered, and the
oracle pins the Sod tube itself."""
import ctypes as C

import numpy as np
import pytest

from parity_util import gpu_config, gpu_hooks, nan_equal_bits

pytestmark = pytest.mark.gpu


def _sod(O, **kw):
    base = dict(N_x=400, N_y=100, hi_y=0.25, Scheme="MHM", Riemann="HLLC", Limiter="minmod", case="sodx",
                Psy_time=0.2)
    base.update(kw)
    return O.Case(**base)


def _solver(case, rows, par):
    import ibtfo_amd as ib
    s = ib.Solver(gpu_config(case))
    ls, init = gpu_hooks(case)
    s.geometry(ls[0], ls[1])
    s.initialize(init[0], init[1])
    s.prepare(case.init_boundary())
    s.set_tile(rows, par)
    return s


def _result(s):
    it, t, dt, rc = s.status()
    d, tl, m = s.log(min(it, 4096))
    return dict(U=s.state(), it=it, t=t, dt=dt, rc=rc, dt_log=d, t_log=tl, mass=m)


def _same(a, b, what):
    assert a["it"] == b["it"], what
    assert a["rc"] == b["rc"], what
    assert nan_equal_bits(a["U"], b["U"]) == 0, what
    assert nan_equal_bits([a["t"], a["dt"]], [b["t"], b["dt"]]) == 0, what
    assert nan_equal_bits(a["dt_log"], b["dt_log"]) == 0, what
    assert nan_equal_bits(a["t_log"], b["t_log"]) == 0, what
    assert np.allclose(a["mass"], b["mass"], rtol=1e-13, atol=0, equal_nan=True), what


def _run(case, rows, par, plan):
    """plan: a list of (steps, graph chunk or 0)"""
    s = _solver(case, rows, par)
    try:
        one = s.one_kernel_steps()
        for n, chunk in plan:
            s.step(n, chunk)
        r = _result(s)
        r["one_kernel"] = one
        return r
    finally:
        s.close()


PLANS = {
    "eager": [(50, 0)] * 8 + [(30, 0)],           # 430 steps asked, the run ends at 357
    "graph100": [(400, 100)],                      # ends inside the fourth replay
    "graph7_tails": [(100, 7), (3, 0), (150, 7), (1, 0), (200, 7)],
}


@pytest.mark.parametrize("plan", sorted(PLANS))
def test_sod_tube_deferred_equals_finish_kernel(oracle_mod, plan):
    O = oracle_mod
    c = _sod(O)
    print("deferred finish, Sod tube 400x100:", plan, flush=True)
    a = _run(c, 3, 1, PLANS[plan])   # row-parallel, 7 x 34 blocks: deferred finish
    b = _run(c, 3, 0, PLANS[plan])   # marching strips: a finish kernel per step
    assert a["one_kernel"] and not b["one_kernel"]
    assert a["it"] == 357
    _same(a, b, plan)
    if plan == "graph100":
        o = O.run_oracle(c, 400, root_mode=1)
        assert o["steps"] == a["it"]
        assert nan_equal_bits(a["U"], o["U_OLD"]) == 0
        assert nan_equal_bits(a["dt_log"], o["dt"]) == 0 and nan_equal_bits(a["t_log"], o["t"]) == 0


@pytest.mark.parametrize("scheme,riemann,lim", [("MUSCL", "HLL", "VanLeer"), ("Godunov", "StegerWarming", "minmod"),
                                                ("MUSCL", "PVRS", "minmod")])
@pytest.mark.parametrize("quirks", [1, 0])
def test_schemes_and_modes(oracle_mod, scheme, riemann, lim, quirks):
    """Reference mode (the HLL y solver's aliased x flux on the strip's first
    row) and corrected mode, on a grid whose strips are ragged in x and y. The
    reference's van Leer macro can drive the tube non-finite, which ends the
    run (dt is NaN): both forms must end it at the same step."""
    O = oracle_mod
    c = _sod(O, N_x=130, N_y=27, Scheme=scheme, Riemann=riemann, Limiter=lim, quirks=quirks, Psy_time=10.0)
    plan = [(40, 8), (5, 0)]
    for rows in (1, 4, 7):
        a = _run(c, rows, 1, plan)
        b = _run(c, rows, 0, plan)
        assert a["one_kernel"] and not b["one_kernel"]
        print("  rows %d: %d steps, status %d" % (rows, a["it"], a["rc"]), flush=True)
        assert a["it"] == 45 or (a["rc"] != 0 and a["it"] > 10)
        _same(a, b, (scheme, riemann, lim, quirks, rows))


def test_callers_dt_and_timing(oracle_mod):
    """ibtfo_step_dt (TimeAdvance's dt, one step per batch) and ibtfo_time_steps
    (events around every step kernel) in the deferred form."""
    import ibtfo_amd as ib
    O = oracle_mod
    c = _sod(O, Psy_time=10.0)
    out = []
    for par in (1, 0):
        s = _solver(c, 3, par)
        try:
            assert s.one_kernel_steps() == bool(par)
            s.step(10)
            for q in range(12):
                dt = s.status()[2] * (0.5 + 0.04 * q)
                assert ib.lib().ibtfo_step_dt(s.h, C.c_double(dt)) == ib.OK
            tot, ker = s.time_steps(20, kernel=True)
            assert tot > 0 and ker > 0
            out.append(_result(s))
        finally:
            s.close()
    assert out[0]["it"] == 10 + 12 + 40
    _same(out[0], out[1], "step_dt + time_steps")


def test_more_blocks_than_the_deferred_form_takes(oracle_mod):
    """1-row strips on 400x100: 700 blocks, beyond the deferred form's 256; the
    row-parallel kernel then runs with a finish kernel, same results."""
    O = oracle_mod
    c = _sod(O)
    a = _run(c, 1, 1, [(120, 100)])
    b = _run(c, 1, 0, [(120, 100)])
    assert not a["one_kernel"] and not b["one_kernel"]
    _same(a, b, "700 blocks")


def test_periodic_ring(oracle_mod):
    """Periodic edges on a grid ragged in x and y (the ring written by the step
    kernel, so nothing runs between a step kernel and its finish): one kernel
    per step, equal to the finish-kernel form and to the oracle."""
    O = oracle_mod
    c = O.Case(N_x=130, N_y=27, lo_x=0.0, lo_y=0.0, hi_x=10.0, hi_y=10.0, Psy_time=1e30, max_iter=10 ** 9,
               CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", case="shu", edge="periodic",
               quirks=0, extra=dict(x0=5.0, y0=5.0, beta=5.0))
    plan = [(40, 8), (5, 0)]
    o = O.run_oracle(c, 45, root_mode=1)
    assert o["steps"] == 45
    for rows in (1, 4, 7):
        print("deferred finish, periodic vortex 130x27, rows", rows, flush=True)
        a = _run(c, rows, 1, plan)
        b = _run(c, rows, 0, plan)
        assert a["one_kernel"] and not b["one_kernel"]
        assert a["it"] == 45
        _same(a, b, ("periodic", rows))
        assert nan_equal_bits(a["U"], o["U_OLD"]) == 0, rows
        assert nan_equal_bits(a["dt_log"], o["dt"]) == 0, rows
