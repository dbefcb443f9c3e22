"""Seeded scratch records (the reference's U_TMP: M*N cells x 7 slots x 4
doubles) for the tests of the record operators (Riemann_solverX/Y,
Conserve2Flux{,X,Y})."""
import numpy as np

RECORD_W = 28
SLOT_OLD, SLOT_LO_FLUX, SLOT_HI_FLUX, SLOT_LO, SLOT_HI = 2, 3, 4, 5, 6


def _states(rng, n, gamma):
    rho = np.exp(rng.uniform(np.log(0.05), np.log(5.0), n))
    p = np.exp(rng.uniform(np.log(0.02), np.log(10.0), n))
    u = rng.uniform(-4.0, 4.0, n)
    v = rng.uniform(-4.0, 4.0, n)
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)], axis=1)


def random_record(nx, ny, ng, seed, gamma=1.4, bad=True):
    """Every slot holds arbitrary values (what a solver leaves alone must stay);
    slots 5 and 6 hold the two states of each face: independent random states
    (subsonic and supersonic faces either way, strong collisions and
    expansions), near-equal pairs on a third of the faces, and with `bad` a few
    non-physical ones (negative pressure, all-zero states) whose NaNs must
    appear in the same places."""
    M, N = nx + 2 * ng, ny + 2 * ng
    n = M * N
    rng = np.random.default_rng(seed)
    rec = rng.uniform(-3.0, 3.0, size=(n, 7, 4))
    lo = _states(rng, n, gamma)
    hi = _states(rng, n, gamma)
    near = rng.random(n) < 0.3
    hi[near] = lo[near] * (1.0 + 1e-3 * rng.standard_normal((int(near.sum()), 4)))
    if bad:
        neg = rng.random(n) < 0.01
        lo[neg, 3] = 0.25 * (lo[neg, 1] ** 2 + lo[neg, 2] ** 2) / lo[neg, 0]  # p < 0
        zero = rng.random(n) < 0.005
        hi[zero] = 0.0
    rec[:, SLOT_LO] = lo
    rec[:, SLOT_HI] = hi
    return np.ascontiguousarray(rec.reshape(-1))
