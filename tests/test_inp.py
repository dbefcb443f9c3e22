"""The .inp grammar: the library's parser against the reference ParamReader's
parse of the four shipped input files and an edge-case file (golden maps written
by tests/golden/make_golden.py from the reference's own ParamReader.cpp)."""
import ctypes as C
import json
import os

import pytest

import ibtfo_amd as ib

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = ["sod.inp", "inclinesod.inp", "vortex.inp", "pnggeo.inp", "edge.inp"]


@pytest.fixture(scope="module")
def parsed():
    with open(os.path.join(GOLD, "inp_parsed.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", FILES)
def test_numbers_match_reference(parsed, name):
    path = os.path.join(GOLD, "inp", name).encode()
    for key, val in parsed[name]["numbers"].items():
        v = C.c_double(-12345.0)
        assert ib.lib().ibtfo_inp_number(path, key.encode("utf-8", "surrogateescape"), C.byref(v)) == 1, key
        assert v.value == val, (name, key, v.value, val)
    v = C.c_double(5.0)
    assert ib.lib().ibtfo_inp_number(path, b"no_such_key", C.byref(v)) == 0 and v.value == 5.0


@pytest.mark.parametrize("name", FILES)
def test_strings_match_reference(parsed, name):
    path = os.path.join(GOLD, "inp", name).encode()
    for key, val in parsed[name]["strings"].items():
        buf = C.create_string_buffer(4096)
        n = ib.lib().ibtfo_inp_string(path, key.encode("utf-8", "surrogateescape"), buf, len(buf))
        got = buf.raw[:max(n, 0)].decode("utf-8", "surrogateescape")
        assert n == len(val.encode("utf-8", "surrogateescape")) and got == val, (name, key, got, val)
    # digit-only values are not strings (ParamReader::is_number)
    if name == "edge.inp":
        buf = C.create_string_buffer(64)
        assert ib.lib().ibtfo_inp_string(path, b"digits", buf, 64) == 0


def test_grammar_edge_cases(parsed):
    n = parsed["edge.inp"]["numbers"]
    assert n["N_x"] == 1.0          # "1e-3" reads as 1
    assert n["Dup"] == 2.5          # last duplicate wins
    assert n["\tTabbed"] == 7.0     # tabs are kept in keys
    assert n["neg"] == -0.25 and n["plus"] == 3.0 and n["hash"] == 12.0 and n["empty"] == 0.0
    assert "#comment line" not in n  # '#' at column 0 skips the line


def test_config_from_shipped_inputs():
    c = ib.config_from_inp(os.path.join(GOLD, "inp", "sod.inp"))
    assert (c.N_x, c.N_y, c.max_iter, c.num_ghost_cell) == (4000, 4000, 100000, 2)
    assert (c.scheme, c.riemann, c.limiter, c.boundary, c.time) == (1, 0, 0, 0, 0)
    assert c.CFL_number == 0.3 and c.Psy_time == 0.1 and c.gamma == 1.4
    c = ib.config_from_inp(os.path.join(GOLD, "inp", "inclinesod.inp"))
    assert (c.N_x, c.Psy_time, c.CFL_number) == (2000, 0.3, 0.3)
    c = ib.config_from_inp(os.path.join(GOLD, "inp", "edge.inp"))
    assert c.scheme == ib.SCHEME["MHM"] and c.riemann == ib.RIEMANN["HLLC"]


def test_slot_names_from_the_makefiles(tmp_path):
    p = tmp_path / "slots.inp"
    # Test/InclineSod/Makefile:1-5 as key=value lines; OIDWBaseGFM names no directory,
    # IDWBaseGFM is the one it means
    p.write_text("Scheme = MUSCL #Lax or MUSCL\nRiemann = HLL\nLimiter = minmod\n"
                 "Boundary = OIDWBaseGFM #Regular or BaseGFM\nTime = SingleStep\n")
    c = ib.config_from_inp(str(p))
    assert (c.scheme, c.riemann, c.limiter, c.boundary, c.time) == (1, 0, 0, ib.BOUNDARY["IDWBaseGFM"], 0)


def test_unknown_slot_is_an_error(tmp_path):
    p = tmp_path / "bad.inp"
    p.write_text("Scheme=WENO\n")
    with pytest.raises(ib.IbtfoError) as e:
        ib.config_from_inp(str(p))
    assert e.value.rc == ib.E_ARG


def test_missing_file_is_io_error(tmp_path):
    with pytest.raises(ib.IbtfoError) as e:
        ib.config_from_inp(str(tmp_path / "none.inp"))
    assert e.value.rc == ib.E_IO
