"""PNG decoding (the mask level set's input) against Pillow, and the CPU
oracle's mask level set against closed-form properties. No GPU needed."""
import io
import os

import numpy as np
import pytest

import ibtfo_amd as ib

PIL = pytest.importorskip("PIL.Image")


def _disc(W, H, cx, cy, r, inside=(0, 0, 0), outside=(255, 255, 255)):
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.zeros((H, W, 4), dtype=np.uint8)
    m = (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 < r * r
    img[...] = outside + (255,)
    img[m] = inside + (255,)
    return img


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P", "1", "I;16"])
def test_decoder_matches_pillow(tmp_path, mode):
    rng = np.random.default_rng(5)
    base = rng.integers(0, 256, size=(37, 53, 4), dtype=np.uint8)
    base[5:20, 10:30] = 0  # a black block (R+G+B = 0) and smooth content for the filters
    base[25:, :, :3] = np.linspace(0, 255, 53, dtype=np.uint8)[None, :, None]
    im = PIL.fromarray(base, "RGBA")
    if mode == "I;16":
        g16 = (base[..., 0].astype(np.uint16) << 8) | base[..., 1]
        im = PIL.fromarray(g16)  # uint16 -> mode I;16
    elif mode != "RGBA":
        im = im.convert(mode)
    path = tmp_path / ("m_%s.png" % mode.replace(";", ""))
    im.save(path)
    got = ib.png_decode(str(path))
    assert got.shape == (37, 53, 4)
    if mode == "I;16":
        raw = np.array(PIL.open(path)).astype(np.uint16)
        want = np.repeat((raw >> 8).astype(np.uint8)[..., None], 3, axis=2)
    else:
        want = np.array(PIL.open(path).convert("RGBA"))[..., :3]  # via RGBA: palette transparency kept apart
    assert np.array_equal(got[..., :3], want)


def test_decoder_rejects_damaged_files(tmp_path):
    p = tmp_path / "ok.png"
    PIL.fromarray(_disc(16, 12, 8, 6, 3)).save(p)
    data = p.read_bytes()
    bad = tmp_path / "bad.png"
    bad.write_bytes(data[: len(data) // 2])
    with pytest.raises(ib.IbtfoError):
        ib.png_decode(str(bad))
    flip = bytearray(data)
    flip[40] ^= 0xFF  # inside a chunk: CRC mismatch
    bad.write_bytes(bytes(flip))
    with pytest.raises(ib.IbtfoError):
        ib.png_decode(str(bad))


def _case(O, nx=40, ny=24):
    return O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=nx / 16.0, hi_y=ny / 16.0, case="uniform")


def test_oracle_mask_phi_disc_properties(oracle_mod):
    """An image of a dark disc on the grid's own resolution: the sign is negative
    inside, |phi| is the distance to the nearest edge pixel, which lies within one
    pixel diagonal of the circle, and cells far from the disc see +distance."""
    O = oracle_mod
    c = _case(O)
    M, N = c.M, c.N
    img = _disc(M, N, cx=20.0, cy=12.0, r=6.0)
    phi = O.mask_phi(c, img).reshape(N, M)
    h = 1.0 / 16
    # cell (i, j) centre; image rows run top-down, the level set bottom-up
    for j in range(2, N - 2):
        for i in range(2, M - 2):
            X, Y = (i - 2 + 0.5) * h, (j - 2 + 0.5) * h
            # disc centre in physical units: pixel (20, 12) from the top -> y from the bottom
            cxp, cyp = 20.0 * h - 2 * h, (N - 12.0) * h - 2 * h
            rr = np.hypot(X - cxp, Y - cyp)
            assert abs(abs(phi[j, i]) - abs(rr - 6 * h)) <= 1.5 * h, (i, j)
            if rr < 6 * h - 1.5 * h:
                assert phi[j, i] < 0
            if rr > 6 * h + 1.5 * h:
                assert phi[j, i] > 0


def test_oracle_mask_phi_no_edges_is_100(oracle_mod):
    O = oracle_mod
    c = _case(O, 12, 8)
    img = np.full((c.N, c.M, 4), 255, dtype=np.uint8)
    assert np.all(O.mask_phi(c, img) == 100.0)
    img[...] = 0
    img[..., 3] = 255
    assert np.all(O.mask_phi(c, img) == -100.0)


def test_oracle_mask_phi_brute_force_numpy(oracle_mod):
    """Against an independent vectorised evaluation on a random mask (edge list,
    nearest-pixel sampling and the 100 cap)."""
    O = oracle_mod
    c = _case(O, 30, 20)
    rng = np.random.default_rng(3)
    W, H = 47, 29
    img = np.zeros((H, W, 4), dtype=np.uint8)
    img[..., :3] = (rng.random((H, W)) < 0.3)[..., None] * 200
    img[..., 3] = 255
    phi = O.mask_phi(c, img)
    v = np.where(img[::-1, :, :3].astype(int).sum(axis=2) > 0, 1.0, -1.0)  # bottom row first
    gx = np.zeros_like(v)
    d1 = np.abs(v[1:-1, 1:-1] - v[1:-1, :-2])
    d2 = np.abs(v[1:-1, 1:-1] - v[:-2, 1:-1])
    gx[1:-1, 1:-1] = np.maximum(d1, d2)
    ys, xs = np.nonzero(gx > 0.1)
    L, Hh = c.L, c.H
    dxp = (L / c.N_x) * (c.N_x + 4) / W
    dyp = (Hh / c.N_y) * (c.N_y + 4) / H
    ex = -2 * (L / c.N_x) + dxp * xs
    ey = -2 * (Hh / c.N_y) + dyp * ys
    dx = L / c.N_x
    X = -2 * dx + dx / 2 + np.arange(c.M) * dx
    Y = -2 * dx + dx / 2 + np.arange(c.N) * dx
    d = np.sqrt((ex[None, None, :] - X[None, :, None]) ** 2 + (ey[None, None, :] - Y[:, None, None]) ** 2).min(axis=2)
    d = np.minimum(d, 100.0)
    xi = ((np.arange(c.M) + 0.5) * W / c.M).astype(int)
    yi = ((np.arange(c.N) + 0.5) * H / c.N).astype(int)
    sgn = np.where(v[yi][:, xi] >= 0, 1.0, -1.0)
    assert np.allclose(phi.reshape(c.N, c.M), sgn * d, rtol=1e-14, atol=1e-15)
