"""Pins of the oracle's functions that have no reference output to compare with
(the corrected arithmetic, the new initialiser and edge policies, VanAlbada):
checks the mathematics fixes, chosen so that a plausible mistake (a dropped
term, a wrong constant, sign or index) fails one of them. No GPU needed."""
import ctypes as C
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from exact_riemann import sample, star_state

G = 1.4


def _cons(r, u, v, p, g=G):
    return np.array([r, r * u, r * v, p / (g - 1) + 0.5 * r * (u * u + v * v)])


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- corrected mode
def _advance(O, c, scheme, U, lim="minmod", rs="HLL", quirks=0, dt=1e-3):
    L = O.lib()
    st = O.setup(c)
    XY = st["XY"]
    T = np.zeros(c.M * c.N * L.or_tmp_width(O.SCHEMES[scheme]))
    U_NEW = U.copy()
    L.or_advance(O.SCHEMES[scheme], O.RIEMANNS[rs], O.LIMITERS[lim], quirks, C.c_double(c.L), C.c_double(c.H),
                 c.N_x, c.N_y, c.num_ghost_cell, C.c_double(c.gamma), C.c_double(dt), _ptr(U), _ptr(T), _ptr(U_NEW),
                 _ptr(XY))
    return U_NEW


def _kinetic_ramp(c, eps, du=0.1):
    """rho = 1, u = du*(i+1) along x, internal energy eps / (gamma-1): the minmod
    slope of the quadratic E is the smaller one-sided difference, so every
    cell's reconstructed state on its right face has internal energy
    eps - 3/8 du^2 (negative for eps < 3/8 du^2) while its left-face state stays
    physical; y is uniform (zero slopes)."""
    U = np.zeros((c.N, c.M, 4))
    for i in range(c.M):
        U[:, i] = _cons(1.0, du * (i + 1), 0.0, (G - 1) * eps)
    return U.reshape(-1)


@pytest.mark.parametrize("rs", ["HLL", "StegerWarming", "PVRS"])
def test_fallback_on_every_face_equals_godunov(oracle_mod, rs):
    """With every x face's reconstructed pair non-physical, the corrected
    MUSCL step falls back to first order on every face: it equals the Godunov
    step bit for bit. With the same profile made physical it does not."""
    O = oracle_mod
    c = O.Case(N_x=24, N_y=6, hi_x=1.0, hi_y=0.25, quirks=0)
    tripped = _kinetic_ramp(c, eps=1e-3)
    a = _advance(O, c, "MUSCL", tripped, rs=rs)
    b = _advance(O, c, "Godunov", tripped, rs=rs)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, tripped)
    ok = _kinetic_ramp(c, eps=2e-2)  # internal energy at the faces stays positive
    a2 = _advance(O, c, "MUSCL", ok, rs=rs)
    b2 = _advance(O, c, "Godunov", ok, rs=rs)
    assert not np.array_equal(a2, b2)
    # the reference arithmetic has no fallback: there the tripped profile differs from Godunov
    a3 = _advance(O, c, "MUSCL", tripped, rs=rs, quirks=1)
    b3 = _advance(O, c, "Godunov", tripped, rs=rs, quirks=1)
    assert not np.array_equal(a3, b3)


def _speeds(O, Ul, Ur, quirks=0):
    S = np.zeros(2)
    O.lib().or_hll_speeds(quirks, 0, C.c_double(G), _ptr(Ul), _ptr(Ur), _ptr(S))
    return S


def test_corrected_speeds_are_the_exact_rarefaction_heads(oracle_mod):
    """Two rarefactions (the exact p* below both pressures, also where the
    linearised p* is negative and clamped at 0): the corrected estimates are the
    exact head speeds u_L - a_L and u_R + a_R."""
    O = oracle_mod
    rng = np.random.default_rng(11)
    clamped = 0
    for _ in range(300):
        rl, rr = rng.uniform(0.2, 3.0, 2)
        pl, pr = rng.uniform(0.1, 3.0, 2)
        al, ar = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
        ul = rng.uniform(-2, 1)
        ur = ul + rng.uniform(0.05, 1.9) * 2 * (al + ar) / (G - 1)  # no vacuum
        p_star, _ = star_state((rl, ul, pl), (rr, ur, pr))
        if not (p_star < pl and p_star < pr):
            continue
        lin = 0.5 * (pl + pr) + 0.5 * (ul - ur) * (0.5 * (rl + rr)) * (0.5 * (al + ar))
        clamped += lin < 0
        S = _speeds(O, _cons(rl, ul, 0.3, pl), _cons(rr, ur, -0.2, pr))
        assert abs(S[0] - (ul - al)) <= 1e-14 * (abs(ul) + al)  # u and a recovered from conserved values
        assert abs(S[1] - (ur + ar)) <= 1e-14 * (abs(ur) + ar)
    assert clamped > 10


def test_corrected_speeds_on_shocks_are_the_reference_form(oracle_mod):
    """Where the linearised p* exceeds a side's pressure (a shock on that side)
    the corrected estimate applies the same shock factor as the reference's
    (HLL/Riemann_solver.cpp:131-168, pinned bit for bit to the reference's own
    build): that side's speed equals the reference form's exactly; a side with
    p* at or below its pressure (a rarefaction) takes the head speed u -/+ a,
    which the reference form does not (its factor is below 1 there)."""
    O = oracle_mod
    rng = np.random.default_rng(12)
    both = one = 0
    for _ in range(600):
        rl, rr = rng.uniform(0.3, 3.0, 2)
        pl, pr = rng.uniform(0.2, 3.0, 2)
        al, ar = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
        ul = rng.uniform(-1, 1)
        ur = ul + rng.uniform(-0.8, 0.3) * (al + ar)
        Ul, Ur = _cons(rl, ul, 0.1, pl), _cons(rr, ur, 0.0, pr)
        pv = 0.5 * (pl + pr) + 0.5 * (ul - ur) * (0.5 * (rl + rr)) * (0.5 * (al + ar))
        S, Sq = _speeds(O, Ul, Ur), _speeds(O, Ul, Ur, quirks=1)
        if pv > pl * (1 + 1e-9):
            assert S[0] == Sq[0]
        elif pv < pl * (1 - 1e-9):
            assert abs(S[0] - (ul - al)) <= 1e-14 * (abs(ul) + al) and Sq[0] > S[0]
        if pv > pr * (1 + 1e-9):
            assert S[1] == Sq[1]
        elif pv < pr * (1 - 1e-9):
            assert abs(S[1] - (ur + ar)) <= 1e-14 * (abs(ur) + ar) and Sq[1] < S[1]
        both += pv > max(pl, pr)
        one += min(pl, pr) < pv < max(pl, pr)
    assert both > 50 and one > 50


@pytest.mark.parametrize("scheme,lim", [("MUSCL", "VanLeer"), ("MHM", "minmod"), ("MUSCL", "minmod")])
def test_corrected_sod_converges_to_exact_solution(oracle_mod, scheme, lim):
    """The corrected arithmetic on the Sod tube (no fallback or floor fires on
    it: equal to the reference form up to the wave-speed estimate) converges to
    the exact Riemann solution (L1 falls by at least 0.7x per grid doubling:
    contacts smear at below first order)."""
    O = oracle_mod
    errs = []
    for nx in (100, 200, 400):
        c = O.Case(N_x=nx, N_y=4, hi_y=4.0 / nx, Scheme=scheme, Limiter=lim, case="sodx", CFL_number=0.5,
                   Psy_time=0.2, quirks=0)
        r = O.run_oracle(c, 100000)
        U = r["U_OLD"].reshape(c.N, c.M, 4)[2, 2:-2]
        x = (np.arange(nx) + 0.5) / nx
        st = O.setup(c)
        dt0 = O.compute_dt(c, st["U_OLD"], st["XY"])
        t_state = dt0 + float(np.sum(r["dt"][:-1]))
        ex = sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), (x - 0.5) / t_state)
        errs.append(np.mean(np.abs(U[:, 0] - ex[:, 0])))
    assert errs[0] < 0.02 and errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1], errs


# ---------------------------------------------------------------- isentropic vortex
def _shu(O, nx, L=10.0, beta=5.0, ng=2):
    c = O.Case(N_x=nx, N_y=nx, hi_x=L, hi_y=L, case="shu", Boundary="Regular", edge="periodic",
               extra=dict(x0=L / 2, y0=L / 2, beta=beta, rho=1.0, u=1.0, v=1.0, p=1.0))
    st = O.setup(c)
    U = st["U_OLD"].reshape(c.N, c.M, 4)
    XY = st["XY"].reshape(c.N, c.M, 6)
    return c, U, XY[..., 0], XY[..., 1]


def test_isentropic_vortex_initialiser(oracle_mod):
    """Shu's isentropic vortex (beta = 5 about (5, 5), free stream (1, 1, 1, 1)):
    p / rho^gamma = 1 everywhere; the velocity perturbation is azimuthal with
    speed beta/(2 pi) r exp((1 - r^2)/2); the radial momentum balance
    dp/dr = rho u_theta^2 / r holds (central differences on a fine grid); the
    far field is the free stream."""
    O = oracle_mod
    c, U, x, y = _shu(O, 2000)
    rho = U[..., 0]
    u, v = U[..., 1] / rho, U[..., 2] / rho
    p = (G - 1) * (U[..., 3] - 0.5 * rho * (u * u + v * v))
    assert np.abs(p / rho ** G - 1).max() < 1e-12
    dx_, dy_ = x - 5.0, y - 5.0
    r = np.hypot(dx_, dy_)
    du, dv = u - 1.0, v - 1.0
    assert np.abs(du * dx_ + dv * dy_).max() < 1e-13
    speed = 5.0 / (2 * math.pi) * r * np.exp(0.5 * (1 - r * r))
    assert np.abs(np.hypot(du, dv) - speed).max() < 1e-13
    # momentum balance along the row through the centre (y = 5 +- h/2)
    j = c.N // 2
    h = x[j, 1] - x[j, 0]
    i = np.arange(3, c.M - 3)
    rr = np.abs(dx_[j, i])
    dpdx = (p[j, i + 1] - p[j, i - 1]) / (2 * h)
    ut2 = du[j, i] ** 2 + dv[j, i] ** 2
    rhs = rho[j, i] * ut2 / r[j, i] * dx_[j, i] / r[j, i] * (rr / r[j, i])
    sel = (rr > 0.3) & (rr < 3.0)
    assert np.abs(dpdx[sel] - rhs[sel]).max() < 2e-5
    far = r > 6.8  # the corners of [0, 10]^2
    assert np.abs(rho[far] - 1).max() < 1e-9 and np.abs(du[far]).max() < 1e-9


# ---------------------------------------------------------------- edge policies
def _ring_case(O, edge):
    c = O.Case(N_x=6, N_y=5, hi_x=1.2, hi_y=1.0, case="uniform", edge=edge,
               extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    return c


@pytest.mark.parametrize("edge", ["inflow", "periodic"])
def test_edge_ring_brute_force(oracle_mod, edge):
    """Every ring cell of a 6 x 5 grid against a brute-force statement of the
    policy: inflow = the inflow state on the columns left of the domain, every
    other ring cell a copy of its nearest interior cell (zero gradient); periodic
    = the interior cell whose centre differs by a multiple of the domain size.
    Interior cells are untouched."""
    O = oracle_mod
    c = _ring_case(O, edge)
    L = O.lib()
    rng = np.random.default_rng(3)
    U_NEW = rng.uniform(0.5, 2.0, size=c.M * c.N * 4)
    U_OLD = np.full_like(U_NEW, -7.0)
    prm = np.array([1.4, 3.0, 0.0, 1.0])
    L.or_global_boundary(O.EDGES[edge], _ptr(prm), c.N_x, c.N_y, c.num_ghost_cell, C.c_double(G), _ptr(U_OLD),
                         _ptr(U_NEW))
    ng, M, N = c.num_ghost_cell, c.M, c.N
    Un, Uo = U_NEW.reshape(N, M, 4), U_OLD.reshape(N, M, 4)
    inner = [(i, j) for j in range(ng, N - ng) for i in range(ng, M - ng)]
    inflow = _cons(1.4, 3.0, 0.0, 1.0)
    for j in range(N):
        for i in range(M):
            if (i, j) in inner:
                assert (Uo[j, i] == -7.0).all()
                continue
            if edge == "inflow":
                if i < ng:
                    assert np.array_equal(Uo[j, i], inflow), (i, j)
                    continue
                d = [(i - a) ** 2 + (j - b) ** 2 for a, b in inner]
                a, b = inner[int(np.argmin(d))]
            else:
                a, b = [(a, b) for a, b in inner if (a - i) % c.N_x == 0 and (b - j) % c.N_y == 0][0]
            assert np.array_equal(Uo[j, i], Un[b, a]), (i, j, a, b)


# ---------------------------------------------------------------- VanAlbada
def test_vanalbada_exact_rationals(oracle_mod):
    """VanAlbada/Slope_limiter.H:4 (its evident intent, the reference's does not
    compile): (r + |r|) b / (1 + |r|), r = (a + 1e-7)/(b + 1e-7), against exact
    rational evaluation (a few ulp for the three roundings); for |a|, |b| >> 1e-7
    it is van Leer's harmonic mean 2ab/(a + b), symmetric in a and b."""
    O = oracle_mod
    Lb = O.lib()
    rng = np.random.default_rng(9)
    vals = list(rng.normal(size=(500, 2))) + [(0.0, 0.0), (1e-8, -3e-8), (2.0, 2.0), (-1.0, -4.0), (3.0, -1.0)]
    for a, b in vals:
        got = Lb.or_limiter(2, 0, a, b, 0.0)
        assert got == Lb.or_limiter(2, 1, a, b, 0.0)  # no reference-mode variant
        A, B = Fr(a + 1e-7), Fr(b + 1e-7)
        if B == 0:
            continue
        r = A / B
        exact = (r + abs(r)) * Fr(b) / (1 + abs(r)) if r > 0 else Fr(0)
        assert abs(Fr(got) - exact) <= Fr(8 * np.spacing(abs(float(exact)) + 1e-300)), (a, b, got, float(exact))
        if a * b > 0 and min(abs(a), abs(b)) > 1e-3:
            off = 4e-6 / min(abs(a), abs(b))  # the 1e-7 offsets perturb r by about 1e-7 / |b|
            assert abs(got - 2 * a * b / (a + b)) <= off * abs(got)
            assert abs(got - Lb.or_limiter(2, 0, b, a, 0.0)) <= off * abs(got)
