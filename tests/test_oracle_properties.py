"""Closed-form and invariant checks of the CPU oracle: what the mathematics fixes
independently of the reference's code (each chosen so that a dropped term, a
wrong sign or index, or a transposed operand fails it)."""
import ctypes as C

import numpy as np
import pytest

from exact_riemann import sample, star_state


def test_exact_riemann_reference_values():
    p, u = star_state((1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    assert abs(p - 0.30313) < 1e-5 and abs(u - 0.92745) < 1e-5


def test_minmod_table(oracle_mod):
    O = oracle_mod
    vals = [-3.0, -2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0, 3.0]
    for a in vals:
        for b in vals:
            want = 0.0 if a * b <= 0 else (a if abs(a) <= abs(b) else b)
            assert O.lib().or_limiter(0, 1, a, b, 0.0) == want, (a, b)


def test_vanleer_macro_defect_and_fixed_form(oracle_mod):
    O = oracle_mod
    L = O.lib()
    # uniform data: a = b = 0 (P = Q = u). The parenthesised formula gives 0 slope;
    # the reference's expanded macro gives (r^2+r)P - Q/(r^2+1) with r = 1: 1.5 u.
    for u in [0.3, 1.0, 2.5]:
        assert L.or_limiter(1, 0, 0.0, u, u) == 0.0
        assert L.or_limiter(1, 1, 0.0, u, u) == 2 * u - u / 2
    # corrected form: (r^2 + r) b / (r^2 + 1) with r = (a+1e-7)/(b+1e-7), the
    # formula the header writes; the oracle evaluates it with one division, which
    # agrees with the two-division evaluation to a few ulp and is exact in rationals
    from fractions import Fraction as Fr
    rng = np.random.default_rng(3)
    vals = list(rng.normal(size=(400, 2))) + [(1e-9, -2e-9), (0.0, 0.0), (3.0, -1e-7), (-1e-7, 2.0)]
    for a, b in vals:
        got = L.or_limiter(1, 0, a, b, 0.0)
        if b + 1e-7 == 0.0:  # r = A/0: the two-division form is NaN, the corrected one 0
            assert got == 0.0
            continue
        r = (a + 1e-7) / (b + 1e-7)
        two_div = (r * r + r) * b / (r * r + 1) if r > 0 else 0.0
        assert abs(got - two_div) <= 8 * np.spacing(abs(two_div)) + 1e-300, (a, b, got, two_div)
        A, B = Fr(a) + Fr(1e-7), Fr(b) + Fr(1e-7)
        exact = Fr(b) * A * (A + B) / (A * A + B * B) if A * B > 0 else Fr(0)
        assert abs(Fr(got) - exact) <= Fr(4 * np.spacing(abs(float(exact)) + 1e-300)), (a, b)


def _cons(r, u, v, p, g=1.4):
    return np.array([r, r * u, r * v, p / (g - 1) + 0.5 * r * (u * u + v * v)])


def test_physical_flux_textbook(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(5)
    for _ in range(200):
        r, p = rng.uniform(0.1, 3, 2)
        u, v = rng.normal(size=2)
        U = _cons(r, u, v, p)
        F, G = np.zeros(4), np.zeros(4)
        O.lib().or_flux(0, C.c_double(1.4), O.ptr(U), O.ptr(F))
        O.lib().or_flux(1, C.c_double(1.4), O.ptr(U), O.ptr(G))
        E = U[3]
        assert np.allclose(F, [r * u, r * u * u + p, r * u * v, (E + p) * u], rtol=1e-13, atol=1e-14)
        assert np.allclose(G, [r * v, r * u * v, r * v * v + p, (E + p) * v], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("rs", [0, 3])  # HLL, Steger-Warming
@pytest.mark.parametrize("ydir", [0, 1])
def test_consistency_F_of_equal_states(oracle_mod, rs, ydir):
    """A two-state flux with equal states is the physical flux (Steger-Warming up
    to its eigenvalue smoothing, delta = 1e-6)."""
    O = oracle_mod
    rng = np.random.default_rng(11)
    for _ in range(200):
        U = _cons(rng.uniform(0.2, 2), rng.normal(), rng.normal(), rng.uniform(0.2, 2))
        F, Fp = np.zeros(4), np.zeros(4)
        O.lib().or_face_flux(rs, ydir, C.c_double(1.4), O.ptr(U), O.ptr(U), O.ptr(F))
        O.lib().or_flux(ydir, C.c_double(1.4), O.ptr(U), O.ptr(Fp))
        tol = 1e-13 if rs == 0 else 1e-9
        assert np.allclose(F, Fp, rtol=tol, atol=tol), (F, Fp)


def test_hll_supersonic_upwinding(oracle_mod):
    O = oracle_mod
    UL, UR = _cons(1.0, 5.0, 0.0, 1.0), _cons(0.5, 4.0, 0.1, 0.4)
    F, FL = np.zeros(4), np.zeros(4)
    O.lib().or_face_flux(0, 0, C.c_double(1.4), O.ptr(UL), O.ptr(UR), O.ptr(F))
    O.lib().or_flux(0, C.c_double(1.4), O.ptr(UL), O.ptr(FL))
    assert np.array_equal(F, FL)


@pytest.mark.parametrize("scheme", ["MUSCL", "MHM", "Godunov", "Lax"])
@pytest.mark.parametrize("rs", ["HLL", "StegerWarming", "PVRS"])
def test_free_stream_preserved_bitwise(oracle_mod, scheme, rs):
    O = oracle_mod
    c = O.Case(N_x=24, N_y=16, hi_y=0.5, Scheme=scheme, Riemann=rs, case="uniform", edge="periodic",
               Psy_time=10.0, extra=dict(rho=1.3, u=0.7, v=-0.4, p=0.9))
    st = O.setup(c)
    U0 = st["U_OLD"].copy()
    r = O.run_oracle(c, 15, state=st)
    assert r["steps"] == 15
    assert np.array_equal(r["U_OLD"], U0)


def test_x_sod_is_y_invariant_and_conservative(oracle_mod):
    O = oracle_mod
    c = O.Case(N_x=200, N_y=8, hi_y=0.04, Scheme="MHM", case="sodx", CFL_number=0.5, Psy_time=0.15)
    r = O.run_oracle(c, 100000)
    U = r["U_OLD"].reshape(c.N, c.M, 4)
    inner = U[2:-2, 2:-2]
    assert all(np.array_equal(inner[j], inner[0]) for j in range(inner.shape[0]))
    # mass conserved until waves reach the walls (shock reaches x = 1 at t ~ 0.28)
    m0 = 0.5 * 1.0 + 0.5 * 0.125
    mass = np.sum(inner[..., 0]) * (1.0 / 200) * (0.04 / 8) / 0.04
    assert abs(mass - m0) < 1e-13


def test_compute_dt_closed_form(oracle_mod):
    O = oracle_mod
    c = O.Case(N_x=20, N_y=10, hi_y=0.5, case="uniform", extra=dict(rho=1.2, u=0.3, v=-0.8, p=2.0))
    st = O.setup(c)
    a = (1.4 * 2.0 / 1.2) ** 0.5
    want = c.CFL_number * (1.0 / 20) / (a + 0.8)
    assert abs(O.compute_dt(c, st["U_OLD"], st["XY"]) - want) <= 1e-15 * want


@pytest.mark.parametrize("nx", [100, 200, 400])
def test_sod_converges_to_exact_solution(oracle_mod, nx):
    """The 400x100 Sod workload's scheme (MHM + minmod + HLL) against the exact Riemann solution:
    L1(rho) falls with resolution (first-order-or-better at the discontinuities)."""
    O = oracle_mod
    c = O.Case(N_x=nx, N_y=4, hi_y=4.0 / nx, Scheme="MHM", case="sodx", CFL_number=0.5, Psy_time=0.2)
    r = O.run_oracle(c, 100000)
    U = r["U_OLD"].reshape(c.N, c.M, 4)[2, 2:-2]
    dx = 1.0 / nx
    x = (np.arange(nx) + 0.5) * dx
    # the reference's clock: each step advances by the dt computed after it, so the
    # state after n steps is at sum of the dt actually used (dt0 .. dt_{n-1})
    dt0 = O.compute_dt(c, O.setup(c)["U_OLD"], O.setup(c)["XY"])
    t_state = dt0 + float(np.sum(r["dt"][:-1]))
    ex = sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), (x - 0.5) / t_state)
    l1 = np.mean(np.abs(U[:, 0] - ex[:, 0]))
    bound = {100: 0.02, 200: 0.011, 400: 0.006}[nx]
    assert l1 < bound, l1


def test_classification_brute_force(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(2)
    c = O.Case(N_x=30, N_y=20, hi_y=20 / 30, case="uniform")
    phi = rng.normal(size=c.M * c.N) + 0.3
    c.phi = phi
    st = O.setup(c)
    t = st["XY"][5::6].reshape(c.N, c.M)
    P = phi.reshape(c.N, c.M)
    ng = 2
    for j in range(c.N):
        for i in range(c.M):
            if not (ng <= i < c.N_x + ng and ng <= j < c.N_y + ng):
                want = 2
            elif P[j, i] >= 0:
                want = 0
            else:
                want = 1
                for k in range(1, ng + 1):
                    for (ii, jj) in [(i - k, j), (i + k, j), (i, j - k), (i, j + k)]:
                        if ng <= ii < c.N_x + ng and ng <= jj < c.N_y + ng and P[jj, ii] >= 0:
                            want = 2
            assert t[j, i] == want, (i, j)


def test_idw_fill_reproduces_wall_parallel_flow(oracle_mod):
    """Uniform flow along an inclined channel: the mirror-point IDW average of a
    uniform field is that field, and its normal velocity is zero, so ghosts take
    the fluid state."""
    O = oracle_mod
    ang = np.deg2rad(30.0)
    u, v = 0.8 * np.cos(ang), 0.8 * np.sin(ang)
    c = O.Case(N_x=60, N_y=60, hi_y=1.0, Boundary="IDWBaseGFM", case="incline", extra=dict(angle=30.0))
    st = O.setup(c)
    U = st["U_OLD"].reshape(-1, 4)
    Uf = _cons(1.1, u, v, 0.7)
    U[:] = Uf
    st["U_NEW"][:] = st["U_OLD"]
    O.lib().or_boundary(1, c.N_x, c.N_y, 2, C.c_double(1.4), O.ptr(st["U_OLD"]), O.ptr(st["U_NEW"]),
                        O.ptr(st["XY"]), O.ptr(st["S"]), 1)
    t = st["XY"][5::6].reshape(c.N, c.M)
    S = st["S"].reshape(c.N, c.M, 8)
    UU = st["U_OLD"].reshape(c.N, c.M, 4)
    n = 0
    for j in range(2, c.N - 2):
        for i in range(2, c.M - 2):
            if t[j, i] == 2 and S[j, i, 4:8].sum() > 0:
                assert np.allclose(UU[j, i], Uf, rtol=1e-9, atol=1e-9), (i, j, UU[j, i])
                n += 1
    assert n > 50


def _density_wave(O, n, time, limiter="VanLeer"):
    """rho = 1 + 0.2 sin(2 pi x), u = 1, p = 1 on the periodic unit interval."""
    c = O.Case(N_x=n, N_y=4, hi_x=1.0, hi_y=4.0 / n, Scheme="MUSCL", Riemann="HLL", Limiter=limiter, Time=time,
               case="uniform", edge="periodic", quirks=0, CFL_number=0.4, Psy_time=0.5,
               extra=dict(rho=1.0, u=1.0, v=0.0, p=1.0))
    st = O.setup(c)
    dx = 1.0 / n
    x = -2 * dx + dx / 2 + np.arange(c.M) * dx
    U = st["U_OLD"].reshape(c.N, c.M, 4)
    rho = 1 + 0.2 * np.sin(2 * np.pi * x)
    U[..., 0] = rho[None, :]
    U[..., 1] = rho[None, :]
    U[..., 2] = 0.0
    U[..., 3] = 1.0 / 0.4 + 0.5 * rho[None, :]
    st["U_NEW"][:] = st["U_OLD"]
    dt0 = O.compute_dt(c, st["U_OLD"], st["XY"])
    r = O.run_oracle(c, 10 ** 6, state=st)
    t = dt0 + float(np.sum(r["dt"][:-1]))
    xi = x[2:-2]
    got = r["U_OLD"].reshape(c.N, c.M, 4)[2, 2:-2, 0]
    return float(np.mean(np.abs(got - (1 + 0.2 * np.sin(2 * np.pi * (xi - t))))))


def test_rk2_heun_is_second_order_and_beats_single_step(oracle_mod):
    """Heun (the reference RK2's intent) with MUSCL + van Leer on a smooth
    periodic density wave: the L1 error falls by more than 3x per refinement
    and is below the single-step update's."""
    O = oracle_mod
    e = [_density_wave(O, n, "RK2") for n in (32, 64, 128)]
    assert e[0] / e[1] > 3.0 and e[1] / e[2] > 3.0, e
    assert _density_wave(O, 128, "SingleStep") > e[2]


def test_rk2_free_stream_preserved_bitwise(oracle_mod):
    O = oracle_mod
    c = O.Case(N_x=24, N_y=16, hi_y=0.5, Scheme="MUSCL", Riemann="HLL", Time="RK2", case="uniform",
               edge="periodic", Psy_time=10.0, extra=dict(rho=1.3, u=0.7, v=-0.4, p=0.9))
    st = O.setup(c)
    U0 = st["U_OLD"].copy()
    r = O.run_oracle(c, 10, state=st)
    assert r["steps"] == 10 and np.array_equal(r["U_OLD"], U0)


def test_pnggeo_shock_initialiser_is_a_normal_shock(oracle_mod):
    """Test/PNGGeo/Initialize.cpp's post-shock state against the textbook
    normal-shock relations (Toro 3.1.3) written independently: density ratio
    (g+1) M^2 / ((g-1) M^2 + 2), pressure ratio 1 + 2g (M^2 - 1) / (g+1), and mass
    flux continuity across a shock moving at speed Ms into still gas; the hook
    takes the shock speed as Ms itself, exact when the upstream sound speed is 1
    (p = 0.71429 = 1/g, the PNGGeo input). Ahead of x = 0.1 the state is the
    upstream one."""
    O = oracle_mod
    g, M = 1.4, 1.3
    c = O.Case(N_x=40, N_y=8, hi_y=0.2, case="pngshock", edge="pnggeo", Psy_time=0.0,
               extra=dict(rho_L=1.0, u_L=0.0, v_L=0.0, p_L=0.71429))
    st = O.setup(c)
    U = st["U_OLD"].reshape(c.N, c.M, 4)
    x = st["XY"][0::6].reshape(c.N, c.M)
    behind, ahead = U[x <= 0.1], U[x > 0.1]
    assert np.all(ahead[:, 0] == 1.0) and np.all(ahead[:, 1] == 0.0) and np.all(ahead[:, 3] == 0.71429 / (g - 1))
    r2, m2, E2 = behind[0, 0], behind[0, 1], behind[0, 3]
    u2 = m2 / r2
    p2 = (g - 1) * (E2 - 0.5 * r2 * u2 * u2)
    assert abs(r2 - (g + 1) * M * M / ((g - 1) * M * M + 2)) <= 1e-15 * r2
    assert abs(p2 / 0.71429 - (1 + 2 * g * (M * M - 1) / (g + 1))) <= 1e-14
    assert abs(1.0 * (0.0 - M) - r2 * (u2 - M)) <= 1e-15
    assert np.all(behind == behind[0]) and np.all(behind[:, 2] == 0.0)
    # the edge hook's inflow state is the same shock's (its constants are these)
    assert np.allclose(O.pnggeo_inflow(g), [r2, u2, 0.0, p2], rtol=1e-15, atol=0)


@pytest.mark.parametrize("scheme,rs", [("MUSCL", "HLL"), ("MUSCL", "PVRS"), ("MHM", "HLL"), ("MHM", "PVRS"),
                                       ("Godunov", "HLL"), ("Lax", "HLL")])
@pytest.mark.parametrize("quirks", [0, 1])
def test_x_mirror_symmetry_bitwise(oracle_mod, scheme, rs, quirks):
    """A state mirror-symmetric about x = 1/2 (u odd, the rest even) on a
    periodic 64x16 grid stays mirror-symmetric bit for bit for 50 steps with
    minmod: the limiter, the HLL/PVRS wave speeds and fluxes and the update are
    odd under x -> -x in floating point too. (Steger-Warming sums its three
    waves in an order that the mirror swaps, and the van Leer limiter's 1e-7
    offsets are not odd, so they agree with their mirror only to rounding.)"""
    from mirror_util import mirror_defects, mirror_state
    O = oracle_mod
    c = O.Case(N_x=64, N_y=16, hi_x=1.0, hi_y=0.25, Scheme=scheme, Riemann=rs, Limiter="minmod", case="uniform",
               quirks=quirks, edge="periodic", Psy_time=1e30, max_iter=10 ** 9, CFL_number=0.5)
    st = O.setup(c)
    st["U_OLD"][:] = mirror_state(64, 16).reshape(-1)
    st["U_NEW"][:] = st["U_OLD"]
    assert mirror_defects(st["U_OLD"], c.M, c.N) == 0
    U0 = st["U_OLD"].copy()
    r = O.run_oracle(c, 50, root_mode=1, state=st)
    assert r["steps"] == 50 and np.isfinite(r["U_OLD"]).all()
    assert np.abs(r["U_OLD"] - U0).max() > 1e-3  # the flow has moved
    assert mirror_defects(r["U_OLD"], c.M, c.N) == 0
