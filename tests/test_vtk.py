"""WriteData's on-disk format (Src/Utils/WriteData.cpp:5-247): the library's
writer produces the reference writer's files byte for byte -- the committed
files the reference wrote for a fixed state (tests/golden/vtk, with a
near-vacuum density, a NaN and a negative pressure planted), and, where the
reference's writer is built in this container (oracle/_ref), a larger state
from both writers side by side."""
import os

import numpy as np
import pytest

import ibtfo_amd as ib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "vtk")


def _cfg(c):
    return ib.config(N_x=c["N_x"], N_y=c["N_y"], lo_x=0.0, lo_y=0.0, hi_x=c["hi_x"], hi_y=c["hi_y"], gamma=1.4)


def test_writer_reproduces_reference_files(tmp_path):
    st = np.load(os.path.join(HERE, "golden", "vtk_state.npz"))
    cfg = _cfg(dict(N_x=28, N_y=20, hi_x=1.0, hi_y=1.0))
    ib.write_vtk_host(cfg, 0, 0.0, st["U"], st["XY"], tmp_path)
    ib.write_vtk_host(cfg, 10, float(st["t"][0]), st["U"], st["XY"], tmp_path)
    for name in ["dat_0_0.vtk", "dat_0_10.vtk", "result.visit"]:
        with open(os.path.join(GOLD, name), "rb") as f:
            want = f.read()
        with open(tmp_path / name, "rb") as f:
            got = f.read()
        assert got == want, name
    # the planted values are in the file as the reference prints them
    text = (tmp_path / "dat_0_10.vtk").read_text()
    assert "nan  \n" in text and "-100.000000  \n" in text and "0.000001  \n" in text


def test_writer_against_reference_writer(oracle_mod, tmp_path):
    O = oracle_mod
    if not os.path.exists(os.path.join(O.REF_DIR, "libref_writedata.so")):
        pytest.skip("oracle/_ref not built")
    c = O.Case(N_x=96, N_y=64, hi_x=1.5, hi_y=1.0, Boundary="BCGFM", case="incline", Psy_time=0.3)
    r = O.run_oracle(c, 25)
    a, b = tmp_path / "ref", tmp_path / "lib"
    a.mkdir()
    b.mkdir()
    t = float(r["t"][r["steps"] - 1])
    for it, tt in [(0, 0.0), (25, t)]:
        O.ref_write_vtk(c, r["U_OLD"], r["XY"], it, tt, a)
        ib.write_vtk_host(_cfg(dict(N_x=96, N_y=64, hi_x=1.5, hi_y=1.0)), it, tt, r["U_OLD"], r["XY"], b)
    for name in ["dat_0_0.vtk", "dat_0_25.vtk", "result.visit"]:
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
