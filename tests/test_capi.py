"""The C-ABI library loads without a GPU and exports every function that
include/ibtfo.h declares; the CPU-only parts of the ABI behave."""
import ctypes as C
import os
import re

import numpy as np

import ibtfo_amd as ib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("ibtfo.h", "ibtfo_ref.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ibtfo_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 25
    L = ib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_signatures_are_exported():
    """The C++ shim exports the reference's free functions with its signatures."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        import pytest
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", ib.LIB_PATH], capture_output=True, text=True).stdout
    for sig in ["Psy_coord(double, double, double, double, int, int, int, double*, int, int)",
                "Level_Set(char*, int, int, int, double*, int, int)",
                "Scheme_Index(int, int, int, double*, double*, int, int)",
                "Boundary(int, int, int, double, double*, double*, double*, double*, int, int)",
                "Global_Boundary(int, int, int, double, double*, double*, int, int)",
                "Advance(double, double, int, int, int, double, double, double*, double*, double*, double*, int, int)",
                "TimeAdvance(double, double, int, int, int, double, double, double*, double*, double*, double*, "
                "double*, int, int)",
                "ComputeDt(double, double, int, int, int, double, double, double*, double*, double*, int, int)",
                "CaculateMass(double, double, int, int, int, double*, double*, double*, int, int)",
                "Riemann_solverX(double, double, int, int, int, double, double*, double*, int, int)",
                "Riemann_solverY(double, double, int, int, int, double, double*, double*, int, int)",
                "Conserve2Flux(int, int, int, double, double*, double*, double*, int, int)",
                "Conserve2FluxX(int, int, int, double, double*, double*, int, int)",
                "Conserve2FluxY(int, int, int, double, double*, double*, int, int)"]:
        assert sig in out, sig


def test_record_operator_arguments_rejected_before_any_device_call():
    L = ib.lib()
    rec = np.zeros(28 * 4)
    u = rec.ctypes.data_as(C.POINTER(C.c_double))
    for args in [(u, None, None, 4, 28), (None, u, None, 4, 28), (u, u, None, -1, 28), (u, u, None, 4, 3)]:
        a, f, g, n, stride = args
        assert L.ibtfo_flux_record_host(C.c_double(1.4), a, f, g, C.c_longlong(n), stride) == ib.E_ARG, args
        assert L.ibtfo_flux_record(C.c_double(1.4), a, f, g, C.c_longlong(n), stride, None) == ib.E_ARG, args
    assert L.ibtfo_riemann_record_host(None, 0, u) == ib.E_ARG
    assert L.ibtfo_riemann_record(None, 1, u) == ib.E_ARG


def test_config_struct_matches_header():
    # field order and size of ibtfo_config as the header declares it
    assert C.sizeof(ib.Config) == 4 * 4 + 7 * 8 + 8 * 4 + 4 * 8


def test_defaults_are_reference_driver_defaults():
    c = ib.config()
    assert (c.N_x, c.N_y, c.max_iter, c.num_ghost_cell) == (500, 200, 100000, 2)
    assert (c.CFL_number, c.Psy_time, c.gamma, c.hi_x, c.hi_y) == (0.3, 1.0, 1.4, 1.0, 1.0)
    assert c.quirks == 1


def test_bad_config_rejected_before_any_device_call():
    for kw in [dict(num_ghost_cell=1), dict(N_x=2), dict(scheme=9), dict(gamma=1.0), dict(Time="RK2", Scheme="Lax"),
               dict(Precision="mixed", quirks=1), dict(Precision="mixed", quirks=0, Scheme="Lax"), dict(precision=7),
               dict(boundary=8)]:
        cfg = ib.config(**kw)
        h = C.c_void_p()
        rc = ib.lib().ibtfo_create(C.byref(cfg), C.byref(h))
        assert rc == ib.E_ARG, kw
        assert ib.lib().ibtfo_last_error()


def test_footprint_scales_with_grid():
    a = ib.lib().ibtfo_footprint(C.byref(ib.config(N_x=1024, N_y=512)))
    b = ib.lib().ibtfo_footprint(C.byref(ib.config(N_x=2048, N_y=1024)))
    assert 3.5 < b / a < 4.5
    big = ib.lib().ibtfo_footprint(C.byref(ib.config(N_x=4096, N_y=2048)))
    assert 0.5e9 < big < 2e9


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """ctypes mirrors of the header's structs have the C compiler's sizes and offsets."""
    import shutil
    import subprocess
    import pytest
    from ibtfo_amd import serving as S
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "ibtfo.h"\nint main(void){printf("%zu %zu %zu %zu '
                   '%zu %zu %zu %zu\\n", sizeof(ibtfo_config), sizeof(ibtfo_batch_config), sizeof(ibtfo_request), '
                   'sizeof(ibtfo_result), sizeof(ibtfo_batch_stats), offsetof(ibtfo_request, arrival), '
                   'offsetof(ibtfo_batch_config, hbm_budget), offsetof(ibtfo_result, t_arrival));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(ib.Config), C.sizeof(S.BatchConfig), C.sizeof(S.Request), C.sizeof(S.Result),
            C.sizeof(S.BatchStats), S.Request.arrival.offset, S.BatchConfig.hbm_budget.offset,
            S.Result.t_arrival.offset]
    assert got == want
