"""Pins the CPU oracle to the reference: bit-exact agreement with the golden
fixtures that tests/golden/make_golden.py wrote from the reference's own
translation units (all NaNs compare equal; NaN payload/sign bits are
compiler-dependent)."""
import json
import os

import numpy as np
import pytest

from parity_util import nan_equal_bits

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def runs():
    with open(os.path.join(GOLD, "runs.json")) as f:
        meta = json.load(f)
    arr = np.load(os.path.join(GOLD, "runs.npz"))
    return meta, arr


def _case(O, m):
    return O.Case(N_x=m["N_x"], N_y=m["N_y"], hi_x=m["hi_x"], hi_y=m["hi_y"], Psy_time=m["Psy_time"],
                  CFL_number=m["CFL_number"], Scheme=m["Scheme"], Riemann=m["Riemann"], Limiter=m["Limiter"],
                  Boundary=m["Boundary"], case=m["case"], edge=m.get("edge", "reference"), extra=m["extra"])


def test_fixture_set_is_complete(runs):
    meta, _ = runs
    assert len(meta) >= 29
    assert {m["Boundary"] for m in meta.values()} >= {"Regular", "IDWBaseGFM", "LinearBaseGFM", "BCGFM", "HBCGFM"}
    assert {m["Scheme"] for m in meta.values()} == {"Lax", "MUSCL", "MHM", "Godunov"}


def test_oracle_reproduces_every_golden_run(oracle_mod, runs):
    O = oracle_mod
    meta, arr = runs
    bad = []
    for name, m in sorted(meta.items()):
        c = _case(O, m)
        r = O.run_oracle(c, m["steps"])
        ok = (r["steps"] == m["steps_taken"]
              and nan_equal_bits(r["U_OLD"], arr[name + "__U"]) == 0
              and np.array_equal(r["XY"][5::6].astype(np.uint8), arr[name + "__type"])
              and nan_equal_bits(r["dt"], arr[name + "__dt"]) == 0
              and nan_equal_bits(r["t"], arr[name + "__t"]) == 0)
        if name + "__S" in arr.files:
            ok = ok and nan_equal_bits(r["S"], arr[name + "__S"]) == 0
            ok = ok and nan_equal_bits(r["XY"].reshape(-1, 6)[:, 2:5], arr[name + "__geom"]) == 0
        if not ok:
            bad.append(name)
    assert not bad, bad


def test_initial_dt_matches(oracle_mod, runs):
    O = oracle_mod
    meta, _ = runs
    for name, m in meta.items():
        c = _case(O, m)
        st = O.setup(c)
        if c.init_boundary():
            O.lib().or_boundary(O.BOUNDS[c.Boundary], c.N_x, c.N_y, 2, O.C.c_double(c.gamma), O.ptr(st["U_OLD"]),
                                O.ptr(st["U_NEW"]), O.ptr(st["XY"]), O.ptr(st["S"]), c.quirks)
        assert O.compute_dt(c, st["U_OLD"], st["XY"]) == m["dt0"], name


def test_limiters_match_reference_macros(oracle_mod):
    O = oracle_mod
    d = np.load(os.path.join(GOLD, "limiters.npz"))
    v = d["inputs"]
    for lim, code in [("minmod", 0), ("VanLeer", 1)]:
        L = np.array([0.5 * O.lib().or_limiter(code, 1, b - a, c, b) for a, b, c, _ in v])
        R = np.array([0.5 * O.lib().or_limiter(code, 1, dd - c, c, b) for _, b, c, dd in v])
        assert nan_equal_bits(L, 0.5 * d[lim + "_L"]) == 0, lim
        assert nan_equal_bits(R, 0.5 * d[lim + "_R"]) == 0, lim
