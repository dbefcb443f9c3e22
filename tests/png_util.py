"""A small PNG encoder for test images (PNG specification, ISO/IEC 15948):
every colour type (0 gray, 2 RGB, 3 palette, 4 gray + alpha, 6 RGBA), bit
depths 1-16, an optional tRNS chunk, rows filtered with a chosen cycle of the
five filter types, and the image data split over several IDAT chunks. Used to
write the mask fixtures the reference's libpng reader decodes."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(tag, data):
    body = tag + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _row_bytes(row, depth):
    """One row of samples (W x channels ints) packed big-endian at `depth` bits."""
    s = np.asarray(row, dtype=np.int64).reshape(-1)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in s)
    if depth == 8:
        return bytes(int(v) for v in s)
    per = 8 // depth
    out = bytearray()
    for k in range(0, len(s), per):
        b = 0
        for q in range(per):
            v = int(s[k + q]) if k + q < len(s) else 0
            b |= v << (8 - depth * (q + 1))
        out.append(b)
    return bytes(out)


def _filter(raw, prev, ftype, bpp):
    out = bytearray(len(raw))
    for k, x in enumerate(raw):
        a = raw[k - bpp] if k >= bpp else 0
        b = prev[k] if prev is not None else 0
        c = prev[k - bpp] if (prev is not None and k >= bpp) else 0
        if ftype == 0:
            p = 0
        elif ftype == 1:
            p = a
        elif ftype == 2:
            p = b
        elif ftype == 3:
            p = (a + b) // 2
        else:
            p = _paeth(a, b, c)
        out[k] = (x - p) & 0xFF
    return bytes([ftype]) + bytes(out)


def encode(samples, color_type, depth, palette=None, trns=None, filters=(0, 1, 2, 3, 4), idat_parts=2):
    """PNG bytes of an image. samples: (H, W) for types 0 and 3, (H, W, C)
    otherwise, as integers below 2**depth. palette: list of (r, g, b); trns:
    bytes of palette alphas (type 3) or the transparent gray / RGB sample."""
    a = np.asarray(samples)
    H, W = a.shape[0], a.shape[1]
    ch = CHANNELS[color_type]
    bpp = max(1, ch * depth // 8)
    raw_rows = [_row_bytes(a[y].reshape(W, ch) if ch > 1 else a[y], depth) for y in range(H)]
    data = bytearray()
    prev = None
    for y, raw in enumerate(raw_rows):
        data += _filter(raw, prev, filters[y % len(filters)], bpp)
        prev = raw
    z = zlib.compress(bytes(data), 9)
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, color_type, 0, 0, 0))
    if palette is not None:
        out += _chunk(b"PLTE", b"".join(bytes(p) for p in palette))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    cut = [len(z) * q // idat_parts for q in range(idat_parts + 1)]
    for q in range(idat_parts):
        out += _chunk(b"IDAT", z[cut[q]:cut[q + 1]])
    out += _chunk(b"IEND", b"")
    return out


def body_mask(W, H, seed):
    """A boolean (H, W) image of a few discs and rectangles (True = body),
    top row first."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    m = np.zeros((H, W), dtype=bool)
    for _ in range(3):
        cx, cy = rng.uniform(0.2, 0.8) * W, rng.uniform(0.25, 0.75) * H
        r = rng.uniform(0.08, 0.2) * min(W, H)
        m |= (xx - cx) ** 2 + (yy - cy) ** 2 < r * r
    x0, y0 = int(rng.uniform(0.1, 0.6) * W), int(rng.uniform(0.1, 0.6) * H)
    m[y0:y0 + max(2, H // 5), x0:x0 + max(2, W // 6)] = True
    return m
