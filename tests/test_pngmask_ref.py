"""The PNG mask path pinned to the reference's own code: fixtures written by
tests/golden/make_pngmask_golden.py from Src/Utils/PngProcess.cpp (libpng
1.6.37) and Test/PNGGeo/Level_Set.cpp built from source (oracle/_ref).

* the library's PNG decoder (ibtfo_png_decode) returns the reference reader's
  bytes for every colour type, bit depth and tRNS form of the fixtures;
* the oracle's mask level set (phi = sign x distance to the nearest edge pixel,
  the nearest-pixel sign sample, normals, classification) equals the
  reference's XYCOORD bit for bit, on grids smaller than, equal to and larger
  than the image;
* where the reference library is built here, the same on seeded random masks.
No GPU needed."""
import json
import os
import shutil

import numpy as np
import pytest

from parity_util import nan_equal_bits

GOLD = os.path.join(os.path.dirname(__file__), "golden", "png")


def _fixtures():
    with open(os.path.join(GOLD, "pngmask.json")) as f:
        meta = json.load(f)
    return meta, np.load(os.path.join(GOLD, "pngmask.npz"))


META, ARR = _fixtures()


def _oracle_xy(O, m, rgba):
    """Psy_coord + the mask level set + normals + classification on the oracle."""
    c = O.Case(N_x=m["N_x"], N_y=m["N_y"], lo_x=m["lo_x"], lo_y=m["lo_y"], hi_x=m["hi_x"], hi_y=m["hi_y"],
               Boundary="IDWBaseGFM", case="uniform", quirks=1, extra=dict(rho=1.0, u=0.0, v=0.0, p=1.0))
    c.phi = O.mask_phi(c, rgba)
    return c, O.setup(c)["XY"]


def _compare_xy(c, got, want):
    """x, y, phi and type of every cell bit for bit; the normals of the interior
    cells (the reference's PNGGeo loops leave the ring's at 0, Src/Geometry's
    fill index 1 and M-2 as well; only ghost cells, which are interior, use them)."""
    g, w = got.reshape(c.N, c.M, 6), want.reshape(c.N, c.M, 6)
    for k in (0, 1, 2, 5):
        assert nan_equal_bits(g[..., k], w[..., k]) == 0, "slot %d" % k
    ng = c.num_ghost_cell
    inner = (slice(ng, c.N - ng), slice(ng, c.M - ng))
    for k in (3, 4):
        assert nan_equal_bits(g[inner + (k,)], w[inner + (k,)]) == 0, "normal slot %d" % k


@pytest.mark.parametrize("name", sorted(META))
def test_decoder_equals_reference_reader(name):
    import ibtfo_amd as ib
    got = ib.png_decode(os.path.join(GOLD, name + ".png"))
    want = ARR[name + "__rgba"]
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", sorted(META))
def test_oracle_mask_level_set_equals_reference(oracle_mod, name):
    O = oracle_mod
    m = META[name]
    c, XY = _oracle_xy(O, m, ARR[name + "__rgba"])
    _compare_xy(c, XY, ARR[name + "__XY"])
    # the fixture has bodies and edges: phi takes distances, not only the sign
    phi = ARR[name + "__XY"][2::6]
    assert (phi < 0).any() and (phi > 0).any() and np.abs(phi).min() < 0.1


def test_fixture_covers_every_colour_type():
    import png_util as P
    types = set()
    for name in META:
        data = open(os.path.join(GOLD, name + ".png"), "rb").read()
        types.add((data[24], data[25]))  # IHDR bit depth, colour type
    assert {t[1] for t in types} == set(P.CHANNELS)
    assert {t[0] for t in types} >= {1, 4, 8, 16}


@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_reference_on_random_masks(oracle_mod, tmp_path, seed):
    """Seeded masks of various image / grid ratios through the reference's
    reader and level set (built here) and through the oracle."""
    import png_util as P
    O = oracle_mod
    if not O.ref_pngmask_available():
        pytest.skip("oracle/_ref/libref_pngmask.so not built")
    rng = np.random.default_rng(100 + seed)
    nx = int(rng.integers(16, 64))
    ny = max(4, nx // 2)
    W, H = int(rng.integers(8, 3 * nx)), int(rng.integers(6, 3 * ny))
    mask = P.body_mask(W, H, seed)
    img = np.where(mask[..., None], 0, rng.integers(1, 256, size=(H, W, 3)))
    (tmp_path / "Demo.png").write_bytes(P.encode(img, 2, 8, filters=(seed % 5,)))
    want = O.ref_mask_level_set(str(tmp_path), nx, ny, 2, (0.0, 0.0, 2.0, 1.0))
    rgba, W2, H2 = O.ref_png_read(str(tmp_path / "Demo.png"))
    m = dict(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0)
    c, got = _oracle_xy(O, m, rgba)
    _compare_xy(c, got, want)
    shutil.rmtree(tmp_path, ignore_errors=True)
