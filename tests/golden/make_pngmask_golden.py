#!/usr/bin/env python3
"""Writes the PNG mask fixtures under tests/golden/png/ from the reference's own
PNG reader and mask level set (oracle/_ref/libref_pngmask.so, built by
oracle/build_ref.py from Src/Utils/PngProcess.cpp and Test/PNGGeo/Level_Set.cpp
under /root/reference, linked against the image's libpng 1.6.37). Run in the
development container only:

    python3 tests/golden/make_pngmask_golden.py

Each case is a PNG file written by tests/png_util.py (colour types 0/2/3/4/6,
bit depths 1-16, tRNS, all five row filters, several IDAT chunks), the RGBA
bytes the reference reader returns for it, and the XYCOORD (x, y, phi, n_x,
n_y, type per cell) the reference's mask level set computes from it as
Demo.png on a grid smaller than, equal to or larger than the image.
"""
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
import png_util as P  # noqa: E402

PNG_DIR = os.path.join(HERE, "png")


def cases():
    """name -> (N_x, N_y, PNG bytes). Body pixels decode to R+G+B = 0."""
    out = {}
    # RGBA8, image = grid with ghosts (44 x 24); a few body pixels transparent,
    # a few background pixels barely non-black
    m = P.body_mask(44, 24, 1)
    a = np.zeros((24, 44, 4), dtype=np.int64)
    a[~m] = (200, 30, 10, 255)
    a[m] = (0, 0, 0, 255)
    a[m & (np.arange(44)[None, :] % 5 == 0)] = (0, 0, 0, 0)
    a[~m & (np.arange(44)[None, :] % 7 == 3)] = (0, 0, 1, 0)
    out["rgba8_equal"] = (40, 20, P.encode(a, 6, 8))
    # RGB8 image smaller than the grid (32 x 17 for 64 x 34)
    m = P.body_mask(32, 17, 2)
    a = np.zeros((17, 32, 3), dtype=np.int64)
    a[~m] = (255, 255, 255)
    a[~m & (np.arange(17)[:, None] % 3 == 1)] = (0, 4, 0)
    out["rgb8_smaller"] = (60, 30, P.encode(a, 2, 8, filters=(4, 3, 2, 1, 0)))
    # 4-bit palette with tRNS, image larger than the grid (100 x 57 for 36 x 20)
    m = P.body_mask(100, 57, 3)
    idx = np.where(m, 0, 1 + (np.arange(100)[None, :] // 7) % 4).astype(np.int64)
    pal = [(0, 0, 0), (255, 0, 0), (0, 120, 0), (1, 0, 0), (9, 9, 200)]
    out["pal4_trns_larger"] = (32, 16, P.encode(idx, 3, 4, palette=pal, trns=bytes([0, 255, 128, 7])))
    # RGBA16: samples below 256 have a zero high byte, which the reader keeps
    # (16-bit samples cut to 8), so 0x00FF is black and 0x0100 is not
    m = P.body_mask(60, 31, 4)
    a = np.zeros((31, 60, 4), dtype=np.int64)
    a[..., 3] = 65535
    a[~m] = (0x0100, 0, 0, 65535)
    a[m] = (0x00FF, 0x0080, 0, 65535)
    a[~m & (np.arange(60)[None, :] % 4 == 0)] = (40000, 5, 60000, 1000)
    out["rgba16_larger"] = (48, 24, P.encode(a, 6, 16, filters=(1, 4)))
    # 1-bit gray, image larger (88 x 48 for 44 x 24)
    m = P.body_mask(88, 48, 5)
    out["gray1_larger"] = (40, 20, P.encode(np.where(m, 0, 1), 0, 1, filters=(2, 0, 3)))
    # gray + alpha 8-bit, image smaller (30 x 15 for 52 x 28)
    m = P.body_mask(30, 15, 6)
    a = np.zeros((15, 30, 2), dtype=np.int64)
    a[..., 0] = np.where(m, 0, 77)
    a[..., 1] = np.where(m, 255, 3)
    out["graya8_smaller"] = (48, 24, P.encode(a, 4, 8))
    # 8-bit gray with a tRNS gray value (alpha only)
    m = P.body_mask(44, 24, 7)
    out["gray8_trns_equal"] = (40, 20, P.encode(np.where(m, 0, 180), 0, 8, trns=bytes([0, 180])))
    # 16-bit RGB with tRNS, image larger
    m = P.body_mask(70, 36, 8)
    a = np.zeros((36, 70, 3), dtype=np.int64)
    a[~m] = (1000, 2000, 65535)
    a[m] = (0, 0, 200)
    out["rgb16_trns_larger"] = (56, 28, P.encode(a, 2, 16, trns=bytes([0, 0, 0, 0, 0, 200]), filters=(3, 4, 0)))
    return out


def main():
    if not O.ref_pngmask_available():
        print("make_pngmask_golden: oracle/_ref/libref_pngmask.so is not built; nothing written")
        return 1
    os.makedirs(PNG_DIR, exist_ok=True)
    meta, arrays = {}, {}
    tmp = tempfile.mkdtemp()
    try:
        for name, (nx, ny, data) in cases().items():
            path = os.path.join(PNG_DIR, name + ".png")
            with open(path, "wb") as f:
                f.write(data)
            rgba, W, H = O.ref_png_read(path)
            shutil.copy(path, os.path.join(tmp, "Demo.png"))
            XY = O.ref_mask_level_set(tmp, nx, ny, 2, (0.0, 0.0, 2.0, 1.0))
            meta[name] = dict(N_x=nx, N_y=ny, ng=2, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, W=W, H=H)
            arrays[name + "__rgba"] = rgba
            arrays[name + "__XY"] = XY
            print("golden png", name, W, H, "grid", nx, ny, "solid", int((XY[5::6] == 1).sum()),
                  "ghost", int((XY[5::6] == 2).sum()), flush=True)
    finally:
        shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(PNG_DIR, "pngmask.npz"), **arrays)
    with open(os.path.join(PNG_DIR, "pngmask.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
