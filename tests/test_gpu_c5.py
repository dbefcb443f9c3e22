"""The request stream at its own scale: 512 seeded immersed-ellipse solves
(seed 1234; grids 64x64 .. 1024x512; Mach 1.5-4 inflow; 500 steps; MUSCL +
van Leer + HLL, IDW ghost fluid) through the batch engine.

* every request gives the same result in the full 512-request mix as alone,
  bit for bit (16 requests, the smallest and the largest among them);
* FP64 requests through the batch equal the CPU oracle bit for bit (the
  device's ellipse level set handed to the oracle as a per-cell field; the
  level set itself is pinned in test_gpu_pngmask.py);
* the mixed-precision form passes the fixed-L1 gate on 16 requests:
  L1(rho_mixed - rho_ref) <= 1.1 L1(rho_fp64 - rho_ref), rho_ref a 2x-refined
  FP64 solve averaged to the grid;
* a request whose state is not finite (NaN inflow) ends with
  IBTFO_E_NONFINITE and leaves every other request bit for bit unchanged."""
import numpy as np
import pytest

from parity_util import nan_equal_bits

pytestmark = pytest.mark.gpu

STEPS = 500
SAMPLE = [35, 152, 425, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12]  # 64x64, 1024x512, 1016x512, ...


def _stream():
    from ibtfo_amd import serving as S
    return S.ellipse_stream(n=512, seed=1234, steps=STEPS)


def _bc(prec, **kw):
    from ibtfo_amd import serving as S
    args = dict(Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM", Precision=prec, quirks=0,
                hbm_budget=64 << 30, max_slots=256, max_step_blocks=65536, max_ghost_blocks=4096,
                max_ring_blocks=8192, rows=16, chunk=50)
    args.update(kw)
    return S.batch_config(**args)


def _alone(bc, item):
    from ibtfo_amd import serving as S
    b = S.Batch(bc)
    try:
        t = b.submit(**dict(item, keep_state=True))
        b.drain()
        return b.result(t)
    finally:
        b.close()


def test_stream_mix_equals_alone():
    from ibtfo_amd import serving as S
    stream = _stream()
    items = S.ellipse_items(stream, precision="mixed")
    keep = set(SAMPLE)
    for i in keep:
        items[i] = dict(items[i], keep_state=True)
    print("request stream: 512 requests in one batch", flush=True)
    b = S.Batch(_bc("mixed"))
    got = {}
    try:
        tickets = {b.submit(**it): i for i, it in enumerate(items)}

        def harvest():
            for t in b.poll():
                i = tickets[t]
                r, U = b.result(t)
                got[i] = (r, U)
                b.release(t)

        while True:
            a, q, _ = b.progress()
            harvest()
            if a == 0 and q == 0:
                break
        harvest()
    finally:
        b.close()
    assert len(got) == 512
    assert all(r.status == 0 and r.iter == STEPS for r, _ in got.values())
    bc = _bc("mixed")
    for i in SAMPLE:
        print("alone: request %d (%dx%d)" % (i, stream[i]["N_x"], stream[i]["N_y"]), flush=True)
        r, U = _alone(bc, items[i])
        rm, Um = got[i]
        assert r.status == 0 and r.iter == STEPS
        assert r.dt == rm.dt and r.now_t == rm.now_t and r.mass == rm.mass
        assert nan_equal_bits(U, Um) == 0, i


@pytest.mark.parametrize("idx", [35, 2, 3])
def test_fp64_requests_equal_oracle(oracle_mod, idx):
    import ibtfo_amd as ib
    from ibtfo_amd import serving as S
    O = oracle_mod
    e = _stream()[idx]
    item = S.ellipse_items([e], precision="FP64")[0]
    print("request %d (%dx%d) FP64 batch vs oracle" % (idx, e["N_x"], e["N_y"]), flush=True)
    r, U = _alone(_bc("FP64"), item)
    s = ib.Solver(item["cfg"])
    try:
        s.geometry(*item["levelset"])
        XY = s.xycoord()
    finally:
        s.close()
    rho, u, v, p = e["inflow"]
    c = O.Case(N_x=e["N_x"], N_y=e["N_y"], lo_x=0.0, lo_y=0.0, hi_x=e["L"], hi_y=1.0, Psy_time=1e30,
               max_iter=STEPS, CFL_number=e["cfl"], Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer",
               Boundary="IDWBaseGFM", case="uniform", edge="inflow", quirks=0, extra=dict(rho=rho, u=u, v=v, p=p))
    c.phi = XY[2::6].copy()
    o = O.run_oracle(c, STEPS, root_mode=1)
    assert o["steps"] == r.iter == STEPS
    assert np.array_equal(o["XY"][5::6], XY[5::6])
    assert nan_equal_bits(U, o["U_OLD"]) == 0
    assert r.dt == o["dt"][-1] and r.now_t == o["t"][-1]


def test_mixed_precision_l1_gate_on_16_requests():
    from ibtfo_amd import serving as S
    stream = _stream()
    items64 = S.ellipse_items(stream, precision="FP64")
    items32 = S.ellipse_items(stream, precision="mixed", keep_state=True)
    bc = _bc("mixed")
    ratios = []
    for i in range(16):
        print("L1 gate: request %d (%dx%d)" % (i, stream[i]["N_x"], stream[i]["N_y"]), flush=True)
        r32, U32 = _alone(bc, items32[i])
        g = S.l1_gate(stream[i], items64[i], U32)
        ratios.append(g["ratio"])
        assert g["l1_fp64"] > 0 and g["ratio"] <= 1.1, (i, g)
    print("L1 ratios mixed / FP64:", ["%.4f" % q for q in ratios], flush=True)


def test_nonfinite_request_leaves_the_others_unchanged():
    from ibtfo_amd import serving as S
    stream = _stream()
    idx = [35, 1, 4, 12, 15, 2, 9, 3, 13, 10, 24, 26]
    items = [dict(S.ellipse_items([stream[i]], precision="FP64")[0], keep_state=True) for i in idx]
    bad = dict(S.ellipse_items([stream[5]], precision="FP64")[0], keep_state=True)
    nan_state = (float("nan"),) + tuple(stream[5]["inflow"][1:])
    bad["init"] = ("uniform", nan_state)

    def run(with_bad):
        b = S.Batch(_bc("FP64", chunk=25))
        try:
            order = list(items)
            if with_bad:
                order.insert(3, bad)
            tickets = [b.submit(**it) for it in order]
            b.drain()
            res = [b.result(t) for t in tickets]
        finally:
            b.close()
        if with_bad:
            return res[:3] + res[4:], res[3]
        return res, None

    print("batch with and without a non-finite request", flush=True)
    clean, _ = run(False)
    mixed, (rb, _) = run(True)
    assert rb.status == S.E_NONFINITE
    for (r1, U1), (r2, U2) in zip(clean, mixed):
        assert r1.status == r2.status == 0
        assert r1.dt == r2.dt and r1.now_t == r2.now_t
        assert nan_equal_bits(U1, U2) == 0
