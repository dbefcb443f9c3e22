"""Batch engine (continuous batching of independent solves): every result is
the same alone and in the mix, bit for bit, and the same as the single-solve
path; admission respects the HBM budget; bad requests are refused."""
import numpy as np
import pytest

from parity_util import nan_equal_bits

pytestmark = pytest.mark.gpu


def _items(prec="FP64", n=5, steps=40, seed=7, time="SingleStep"):
    from ibtfo_amd import serving as S
    st = S.ellipse_stream(n=n, seed=seed, steps=steps)
    # small grids for the test: scale each request to at most 256 x 128
    for s in st:
        f = max(1, max(s["N_x"] // 256, s["N_y"] // 128))
        s["N_x"], s["N_y"] = max(64, s["N_x"] // f // 8 * 8), max(64, s["N_y"] // f // 8 * 8)
        s["L"] = s["N_x"] / s["N_y"]
        e = s["ellipse"]
        s["ellipse"] = (0.3 * s["L"], 0.5, e[2], e[3], e[4])
    return st, S.ellipse_items(st, precision=prec, keep_state=True, Time=time)


def _bc(prec, time="SingleStep", **kw):
    from ibtfo_amd import serving as S
    args = dict(Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM", Time=time,
                Precision=prec, quirks=0, hbm_budget=1 << 30, max_slots=16, max_step_blocks=4096,
                max_ghost_blocks=1024, max_ring_blocks=1024, rows=16, chunk=8)
    args.update(kw)
    return S.batch_config(**args)


def _run_alone(item, prec, time="SingleStep"):
    from ibtfo_amd import serving as S
    b = S.Batch(_bc(prec, time))
    t = b.submit(**item)
    b.drain()
    r, U = b.result(t)
    b.close()
    return r, U


@pytest.mark.parametrize("prec", ["FP64", "mixed"])
def test_alone_equals_mix_with_staggered_arrivals(prec):
    from ibtfo_amd import serving as S
    st, items = _items(prec)
    alone = [_run_alone(it, prec) for it in items]
    b = S.Batch(_bc(prec))
    tickets = {}
    # two requests first; the rest arrive while those are on the GPU
    for i in (0, 1):
        tickets[b.submit(**items[i])] = i
    b.progress()
    b.progress()
    for i in range(2, len(items)):
        tickets[b.submit(**items[i])] = i
        b.progress()
    order = b.drain()
    assert b.stats().finished == len(items)
    for t, i in tickets.items():
        r, U = b.result(t)
        ra, Ua = alone[i]
        assert r.status == 0 and r.iter == items[i]["cfg"].max_iter
        assert nan_equal_bits(U, Ua) == 0, i
        assert r.dt == ra.dt and r.now_t == ra.now_t
    b.close()


def test_many_queued_at_once_equal_alone():
    """24 requests queued at once on 16 slots: a burst is admitted, the rest
    while chunks run and as slots free up; every result equals its run alone
    bit for bit."""
    from ibtfo_amd import serving as S
    st, items = _items("FP64", n=24, seed=11)
    alone = [_run_alone(it, "FP64") for it in items]
    b = S.Batch(_bc("FP64"))
    tickets = {b.submit(**it): i for i, it in enumerate(items)}
    b.drain()
    assert b.stats().finished == len(items)
    for t, i in tickets.items():
        r, U = b.result(t)
        ra, Ua = alone[i]
        assert r.status == 0 and r.iter == items[i]["cfg"].max_iter
        assert nan_equal_bits(U, Ua) == 0, i
        assert r.dt == ra.dt and r.now_t == ra.now_t
    b.close()


def test_batch_equals_single_solve_path():
    import ibtfo_amd as ib
    st, items = _items("FP64", n=3)
    for it in items:
        r, U = _run_alone(it, "FP64")
        s = ib.Solver(it["cfg"])
        s.geometry(*it["levelset"])
        s.initialize(*it["init"])
        s.prepare(True)
        s.step(it["cfg"].max_iter)
        i, t, dt, rc = s.status()
        assert i == r.iter and t == r.now_t and dt == r.dt
        assert nan_equal_bits(s.state(), U) == 0
        s.close()


def test_rk2_in_batch_equals_single_solve_path():
    import ibtfo_amd as ib
    st, items = _items("FP64", n=2, time="RK2")
    for it in items:
        r, U = _run_alone(it, "FP64", time="RK2")
        s = ib.Solver(it["cfg"])
        s.geometry(*it["levelset"])
        s.initialize(*it["init"])
        s.prepare(True)
        s.step(it["cfg"].max_iter)
        assert nan_equal_bits(s.state(), U) == 0
        s.close()


def test_admission_by_hbm_budget():
    """A budget that holds about two of the solves: all finish, never more than the
    budget is resident, results equal the unconstrained run."""
    import ctypes as C
    import ibtfo_amd as ib
    from ibtfo_amd import serving as S
    st, items = _items("FP64", n=6, steps=24)
    foot = [ib.lib().ibtfo_footprint(C.byref(it["cfg"])) for it in items]
    budget = int(2.2 * max(foot))
    b = S.Batch(_bc("FP64", hbm_budget=budget))
    tickets = {b.submit(**it): i for i, it in enumerate(items)}
    peak_active, peak_used = 0, 0
    while True:
        a, q, f = b.progress()
        s = b.stats()
        peak_active = max(peak_active, a)
        peak_used = max(peak_used, s.arena_used)
        assert s.arena_used <= s.arena_capacity
        b.poll()
        if a == 0 and q == 0:
            break
    assert 1 <= peak_active < len(items)
    big = S.Batch(_bc("FP64"))
    for t, i in tickets.items():
        r, U = b.result(t)
        t2 = big.submit(**items[i])
        big.drain()
        r2, U2 = big.result(t2)
        assert nan_equal_bits(U, U2) == 0
    b.close()
    big.close()


def test_refused_requests():
    import ibtfo_amd as ib
    from ibtfo_amd import serving as S
    st, items = _items("FP64", n=1)
    b = S.Batch(_bc("FP64", hbm_budget=1 << 24))
    huge = ib.config(N_x=4096, N_y=2048, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM",
                     quirks=0, Edge="inflow")
    with pytest.raises(ib.IbtfoError):
        b.submit(cfg=huge, levelset=("none", ()), init=("uniform", (1.4, 3.0, 0.0, 1.0)))
    wrong = items[0]["cfg"]
    wrong.limiter = ib.LIMITER["minmod"]
    with pytest.raises(ib.IbtfoError):
        b.submit(cfg=wrong, levelset=items[0]["levelset"], init=items[0]["init"])
    b.close()
    # batch selections the engine does not run: refused at creation, with a message
    for bad in ["CutCell", "SimpleBound"]:
        with pytest.raises(ib.IbtfoError, match="batch engine"):
            S.Batch(_bc("FP64", Boundary=bad))


def test_request_over_the_grid_capacity_is_refused_or_retired():
    """A request whose step strips exceed the batch's grid capacity is refused
    at submit; one whose ghost list exceeds the ghost-block capacity (known only
    after its geometry) is retired with IBTFO_E_NOMEM once the batch is empty,
    and the requests behind it still run."""
    import ibtfo_amd as ib
    from ibtfo_amd import serving as S
    st, items = _items("FP64", n=3, steps=8)
    b = S.Batch(_bc("FP64", max_step_blocks=8))
    with pytest.raises(ib.IbtfoError, match="step blocks"):
        b.submit(**items[0])
    b.close()
    ghosts = []
    for it in items:
        s = ib.Solver(it["cfg"])
        s.geometry(*it["levelset"])
        ghosts.append(s.num_ghosts())
        s.close()
    # capacity for the smallest ghost list only
    small = min(ghosts)
    cap = (small + 127) // 128
    big = [i for i, g in enumerate(ghosts) if (g + 127) // 128 > cap]
    assert big, ghosts
    b = S.Batch(_bc("FP64", max_ghost_blocks=cap))
    tickets = {b.submit(**it): i for i, it in enumerate(items)}
    b.drain(max_rounds=10000)
    assert b.stats().finished == len(items)
    for t, i in tickets.items():
        r, _ = b.result(t, state=False)
        if i in big:
            assert r.status == ib.E_NOMEM and r.iter == 0
        else:
            assert r.status == 0 and r.iter == items[i]["cfg"].max_iter
    b.close()


@pytest.mark.parametrize("prec,time", [("FP64", "SingleStep"), ("mixed", "SingleStep"), ("FP64", "RK2")])
def test_persistent_kernel_equals_per_phase_launches(prec, time):
    """The persistent batched step (one launch per chunk: every resident solve's
    steps, phases separated by grid barriers) gives the results of the
    one-launch-per-phase sequence bit for bit, with requests arriving while
    chunks run and finishing at different times."""
    from ibtfo_amd import serving as S
    st, items = _items(prec, n=10, steps=36, seed=5, time=time)

    def run(persistent):
        b = S.Batch(_bc(prec, time, persistent=persistent, chunk=8))
        try:
            tickets = {}
            for i in range(4):
                tickets[b.submit(**items[i])] = i
            b.progress()
            for i in range(4, len(items)):
                tickets[b.submit(**items[i])] = i
                b.progress()
            b.drain()
            assert b.stats().finished == len(items)
            return {i: b.result(t) for t, i in tickets.items()}
        finally:
            b.close()

    print("persistent vs per-phase batch, %s %s" % (prec, time), flush=True)
    a, p = run(0), run(1)
    for i in range(len(items)):
        (ra, Ua), (rp, Up) = a[i], p[i]
        assert ra.status == rp.status == 0 and ra.iter == rp.iter == items[i]["cfg"].max_iter
        assert ra.dt == rp.dt and ra.now_t == rp.now_t
        assert nan_equal_bits(Ua, Up) == 0, i


def test_batch_autotune_keeps_results():
    """The batch autotuner times the sample resident together for each strip
    height and launch mode, discards it and keeps the fastest; the batch then
    serves requests whose states equal their runs with the default tile, bit
    for bit (the strip shape changes only the order of the mass diagnostic's
    partial sums, so mass is compared to rounding)."""
    from ibtfo_amd import serving as S
    st, items = _items("FP64", n=6, steps=24, seed=9)
    b = S.Batch(_bc("FP64"))
    try:
        print("batch autotune", flush=True)
        rows, per, ms = b.autotune(items[:4], probe_chunks=1)
        assert rows in (4, 8, 16, 32) and per in (0, 1) and ms > 0
        assert b.stats().active == 0 and b.stats().arena_used == 0
        tickets = {b.submit(**it): i for i, it in enumerate(items)}
        b.drain()
        got = {i: b.result(t) for t, i in tickets.items()}
    finally:
        b.close()
    for i, it in enumerate(items):
        r0, U0 = _run_alone(it, "FP64")
        r, U = got[i]
        assert r.status == 0 and r.iter == r0.iter and r.dt == r0.dt and r.now_t == r0.now_t
        assert abs(r.mass - r0.mass) <= 1e-12 * abs(r0.mass)
        assert nan_equal_bits(U, U0) == 0, i
