"""Synthetic problem definitions used by bench.py, the tests and the serving
demo (seeded, so every run sees the same geometry)."""
import math

import numpy as np

SHAPE_STRIDE = 21


def bodies_shapes(seed=20261202, H=1.0):
    """8-20 bodies (circles or convex polygons with 3-8 vertices), radius
    U(0.02, 0.1)*H, centres in [0.2, 1.9] x [0.1, 0.9]; returns the flat
    IBTFO_LS_SHAPES parameter list."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(8, 21))
    prm = [float(n)]
    for _ in range(n):
        rec = [0.0] * SHAPE_STRIDE
        r = float(rng.uniform(0.02, 0.1) * H)
        cx, cy = float(rng.uniform(0.2, 1.9)), float(rng.uniform(0.1, 0.9))
        if rng.random() < 0.5:
            rec[0], rec[1], rec[2], rec[3] = 0.0, cx, cy, r
        else:
            nv = int(rng.integers(3, 9))
            ang = np.sort(rng.uniform(0, 2 * math.pi, nv))
            rec[0], rec[1], rec[2], rec[3], rec[4] = 1.0, cx, cy, r, float(nv)
            for e in range(nv):
                rec[5 + 2 * e] = cx + r * math.cos(ang[e])
                rec[6 + 2 * e] = cy + r * math.sin(ang[e])
        prm += rec
    return prm


def shapes_phi(prm, x, y):
    """numpy evaluation of the IBTFO_LS_SHAPES signed distance (same formulas)."""
    n = int(prm[0])
    best = np.full(np.broadcast(x, y).shape, 1e300)
    for k in range(n):
        sh = prm[1 + k * SHAPE_STRIDE: 1 + (k + 1) * SHAPE_STRIDE]
        if sh[0] == 0.0:
            d = np.sqrt((x - sh[1]) * (x - sh[1]) + (y - sh[2]) * (y - sh[2])) - sh[3]
        else:
            nv = int(sh[4])
            dmin = np.full_like(best, 1e300)
            inside = np.ones_like(best, dtype=bool)
            for e in range(nv):
                ax, ay = sh[5 + 2 * e], sh[6 + 2 * e]
                bx, by = sh[5 + 2 * ((e + 1) % nv)], sh[6 + 2 * ((e + 1) % nv)]
                ex, ey = bx - ax, by - ay
                px, py = x - ax, y - ay
                t = np.clip((px * ex + py * ey) / (ex * ex + ey * ey), 0.0, 1.0)
                qx, qy = px - t * ex, py - t * ey
                dmin = np.minimum(dmin, qx * qx + qy * qy)
                inside &= ~(ex * py - ey * px < 0.0)
            d = np.where(inside, -np.sqrt(dmin), np.sqrt(dmin))
        best = np.minimum(best, d)
    return best


def bodies_mask(N_x, N_y, ng=2, L=2.0, H=1.0, seed=20261202):
    """The bodies of bodies_shapes(seed) as a binary mask image: RGBA, one pixel
    per cell including the ghost ring ((N_x + 2 ng) x (N_y + 2 ng) pixels, top
    row first), black (0, 0, 0, 255) where the pixel's point lies inside a body,
    white elsewhere. Pixel (x, y), counted from the bottom row, stands at
    (-ng dx + x dx, -ng dy + y dy), where the mask level set places it
    (Test/PNGGeo/Level_Set.cpp:82-95)."""
    dx, dy = L / N_x, H / N_y
    M, N = N_x + 2 * ng, N_y + 2 * ng
    prm = bodies_shapes(seed, H)
    img = np.full((N, M, 4), 255, np.uint8)
    px = -ng * dx + np.arange(M) * dx
    for y0 in range(0, N, 256):  # row blocks keep the temporaries small at 4096 x 2048
        py = -ng * dy + np.arange(y0, min(N, y0 + 256)) * dy
        X, Y = np.meshgrid(px, py)
        solid = shapes_phi(prm, X, Y) < 0.0
        img[N - 1 - y0 - np.arange(len(py))[:, None], np.arange(M)[None, :], :3] = np.where(solid, 0, 255)[..., None]
    return img


def write_png(path, rgba):
    """An (H, W, 4) uint8 image as an 8-bit RGBA PNG (rows unfiltered, one
    zlib-compressed IDAT chunk), the form the mask level set's reader takes."""
    import struct
    import zlib
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = img.shape[:2]
    raw = np.zeros((h, 1 + 4 * w), np.uint8)  # filter byte 0 ("None") before each row
    raw[:, 1:] = img.reshape(h, 4 * w)

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)))
        f.write(chunk(b"IEND", b""))


def shu_exact(x, y, t, gam=1.4, beta=5.0, L=10.0):
    """Density of the isentropic vortex (strength beta, centred at (5, 5) at t = 0)
    advected by the free stream (1, 1) on the periodic [0, L]^2."""
    xc = (5.0 + t) % L
    yc = (5.0 + t) % L
    dxv = (x - xc + L / 2) % L - L / 2
    dyv = (y - yc + L / 2) % L - L / 2
    r2 = dxv * dxv + dyv * dyv
    T = 1 - (gam - 1) * beta * beta / (8 * gam * math.pi ** 2) * np.exp(1 - r2)
    return T ** (1 / (gam - 1))


# Mach 3 inflow (rho, u, v, p) of the 4096x2048 immersed-body workload
MACH3_INFLOW = (1.4, 3.0, 0.0, 1.0)


def mach3_bodies_config(ib, N_x=4096, N_y=2048, Riemann="HLL", Limiter="VanLeer", Boundary="CutCell", quirks=0,
                        max_iter=10 ** 9, CFL_number=0.5):
    return ib.config(N_x=N_x, N_y=N_y, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=max_iter,
                     CFL_number=CFL_number, gamma=1.4, Scheme="MUSCL", Riemann=Riemann, Limiter=Limiter,
                     Boundary=Boundary, Edge="inflow", quirks=quirks, edge_state=MACH3_INFLOW)
