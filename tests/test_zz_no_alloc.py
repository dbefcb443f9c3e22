"""ibtfo_compute_dt_noalloc: ComputeDt + CaculateMass with the partials in the
library's fixed device array (no allocation, no synchronisation) equals
ibtfo_compute_dt + ibtfo_mass bit for bit, on two calls in a row (the second on
a smaller solve), leaves the solve's own run unchanged, and refuses solves
larger than its array holds."""
import ctypes as C

import numpy as np
import pytest

from parity_util import nan_equal_bits

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory for the test's output through the HIP runtime the library
    is linked against (libamdhip64.so.7): allocated and freed by the test, never
    by the entry point under test."""

    def __init__(self, n):
        import ibtfo_amd as ib
        ib.lib()  # loads the runtime
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0
        host = np.full(n, -1.0)
        assert self.hip.hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0  # H2D

    def get(self):
        host = np.zeros(self.n)
        assert self.hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0  # D2H
        return host

    def free(self):
        if self.p.value:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _solve(nx, ny, bound, steps):
    import ibtfo_amd as ib
    cfg = ib.config(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=10 ** 9,
                    CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary=bound,
                    Edge="inflow", quirks=0, edge_state=(1.4, 3.0, 0.0, 1.0))
    s = ib.Solver(cfg)
    s.geometry("ellipse", (0.5, 0.5, 0.15, 0.06, 30.0))
    s.initialize("uniform", (1.4, 3.0, 0.0, 1.0))
    s.prepare(True)
    s.step(steps)
    return s


def _noalloc(s, out):
    import ibtfo_amd as ib
    return ib.lib().ibtfo_compute_dt_noalloc(s.h, out.p)


def test_noalloc_equals_compute_dt_and_mass():
    import ibtfo_amd as ib
    out = DeviceBuffer(3)
    try:
        for nx, ny, bound, steps in [(512, 256, "IDWBaseGFM", 12), (200, 100, "CutCell", 7)]:
            print("compute_dt_noalloc %dx%d %s" % (nx, ny, bound), flush=True)
            s = _solve(nx, ny, bound, steps)
            try:
                assert _noalloc(s, out) == ib.OK, ib.lib().ibtfo_last_error()
                s.status()  # waits for the solve's stream
                got = out.get()
                dt, m = s.compute_dt(), s.mass()
                assert got[0] == dt and got[1] == m and got[2] == 0.0, (got, dt, m)
            finally:
                s.close()
    finally:
        out.free()


def test_noalloc_leaves_the_run_unchanged():
    import ibtfo_amd as ib
    out = DeviceBuffer(3)
    a, b = _solve(256, 128, "IDWBaseGFM", 5), _solve(256, 128, "IDWBaseGFM", 5)
    try:
        for _ in range(3):
            assert _noalloc(a, out) == ib.OK
            a.step(4)
            b.step(4)
        assert a.status()[:3] == b.status()[:3]
        assert nan_equal_bits(a.state(), b.state()) == 0
        assert a.log(17)[0].tolist() == b.log(17)[0].tolist()
    finally:
        a.close()
        b.close()
        out.free()


def test_noalloc_refuses_grids_beyond_its_array():
    import ibtfo_amd as ib
    out = DeviceBuffer(3)
    s = _solve(1100, 1000, "Regular", 0)
    try:
        assert _noalloc(s, out) == ib.E_ARG
        assert b"interior cells" in ib.lib().ibtfo_last_error()
        assert (out.get() == -1.0).all()  # nothing launched
    finally:
        s.close()
        out.free()
