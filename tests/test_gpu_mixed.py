"""The mixed-precision form (FP32 reconstruction and fluxes, FP64 conserved
variables, update, ghost fill and reductions) against FP64.

Bars: no bit-exactness (the form changes the arithmetic); the L1 density error
is stated against the FP64 oracle, and on the isentropic vortex the mixed
form's L1 error against the exact solution is at most 1.1x the FP64 path's."""
import math

import numpy as np
import pytest

from parity_util import gpu_run

pytestmark = pytest.mark.gpu


def _bodies(O, nx=256, ny=128):
    from ibtfo_amd import workloads as W
    c = O.Case(N_x=nx, N_y=ny, lo_x=0.0, lo_y=0.0, hi_x=2.0, hi_y=1.0, Psy_time=1e30, max_iter=10 ** 9,
               CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="IDWBaseGFM",
               case="uniform", edge="inflow", quirks=0, extra=dict(rho=1.4, u=3.0, v=0.0, p=1.0))
    dx = 2.0 / nx
    Li = -dx * 2 + dx / 2
    x = (Li + np.arange(c.M) * dx)[None, :].repeat(c.N, 0)
    y = (Li + np.arange(c.N) * dx)[:, None].repeat(c.M, 1)
    c.phi = W.shapes_phi(W.bodies_shapes(), x, y).reshape(-1)
    return c


def _fluid_l1(U, ref, types, M, N):
    t = types.reshape(N, M)
    r = U.reshape(N, M, 4)[..., 0]
    q = ref.reshape(N, M, 4)[..., 0]
    m = t == 0
    return float(np.mean(np.abs(r[m] - q[m]))), float(np.mean(np.abs(q[m])))


@pytest.mark.parametrize("riemann,boundary", [("HLL", "IDWBaseGFM"), ("StegerWarming", "IDWBaseGFM"),
                                               ("HLL", "CutCell"), ("StegerWarming", "CutCell")])
def test_mixed_bodies_l1_against_fp64_oracle(oracle_mod, riemann, boundary):
    """Mixed form on the Mach 3 bodies (ghost fluid and cut cells; the cut-cell
    apertures, volume fractions, wall force and merging stay FP64): L1 density
    difference to the FP64 oracle after 60 steps <= 1e-4 of the mean density."""
    O = oracle_mod
    c = _bodies(O)
    c.Riemann = riemann
    c.Boundary = boundary
    o = O.run_oracle(c, 60, root_mode=1)
    import ibtfo_amd as ib
    from parity_util import gpu_config
    from ibtfo_amd import workloads as W
    s = ib.Solver(gpu_config(c, precision=ib.PRECISION["mixed"]))
    s.geometry("field", (), phi=c.phi)
    s.initialize("uniform", W.MACH3_INFLOW)
    s.prepare(True)
    s.step(60)
    it, t, dt, rc = s.status()
    assert it == 60 and rc == 0
    U = s.state()
    assert np.isfinite(U).all()
    l1, scale = _fluid_l1(U, o["U_OLD"], s.xycoord()[5::6], c.M, c.N)
    print("mixed %s %s: L1(rho - rho_fp64 oracle) = %.3e (mean rho %.3f), dt %.6e vs %.6e"
          % (riemann, boundary, l1, scale, dt, o["dt"][-1]), flush=True)
    assert l1 <= 1e-4 * scale
    assert abs(dt - o["dt"][-1]) <= 1e-4 * o["dt"][-1]


def shu_exact(x, y, t, gam=1.4, beta=5.0, L=10.0):
    """Isentropic vortex advected by (1, 1) on the periodic [0, L]^2: density."""
    xc = (5.0 + t) % L
    yc = (5.0 + t) % L
    dxv = (x - xc + L / 2) % L - L / 2
    dyv = (y - yc + L / 2) % L - L / 2
    r2 = dxv * dxv + dyv * dyv
    T = 1 - (gam - 1) * beta * beta / (8 * gam * math.pi ** 2) * np.exp(1 - r2)
    return T ** (1 / (gam - 1))


def _vortex_run(prec, n, t_end):
    import ibtfo_amd as ib
    cfg = ib.config(N_x=n, N_y=n, lo_x=0.0, lo_y=0.0, hi_x=10.0, hi_y=10.0, Psy_time=t_end, max_iter=10 ** 9,
                    CFL_number=0.5, Scheme="MUSCL", Riemann="HLL", Limiter="VanLeer", Boundary="Regular",
                    Time="RK2", Edge="periodic", quirks=0, precision=ib.PRECISION[prec])
    s = ib.Solver(cfg)
    s.geometry("none", ())
    s.initialize("shu", (5.0, 5.0, 5.0, 1.0, 1.0, 1.0, 1.0))
    s.prepare(False)
    s.run(graph_chunk=100)
    it, t, dt, rc = s.status()
    U = s.state().reshape(n + 4, n + 4, 4)[2:-2, 2:-2, 0]
    s.close()
    return U, it, t, rc


@pytest.mark.parametrize("n,t_end", [(128, 2.0), (1024, 10.0)])
def test_mixed_vortex_l1_gate(n, t_end):
    """Isentropic vortex (strength 5, free stream (1,1,1,1), periodic [0,10]^2,
    MUSCL + van Leer + HLL with Heun's RK2; the single-step update amplifies
    perturbations of smooth flow over thousands of steps):
    L1(rho_mixed - rho_exact) <= 1.1 L1(rho_fp64 - rho_exact)."""
    U64, it64, t64, rc64 = _vortex_run("fp64", n, t_end)
    U32, it32, t32, rc32 = _vortex_run("mixed", n, t_end)
    assert rc64 == 0 and rc32 == 0
    h = 10.0 / n
    xc = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc)
    e64 = float(np.mean(np.abs(U64 - shu_exact(X, Y, t64))))
    e32 = float(np.mean(np.abs(U32 - shu_exact(X, Y, t32))))
    d = float(np.mean(np.abs(U32 - U64)))
    print("vortex %d^2 t=%.3f/%.3f steps %d/%d: L1 fp64 %.4e mixed %.4e (ratio %.4f), L1(mixed - fp64) %.3e"
          % (n, t64, t32, it64, it32, e64, e32, e32 / e64, d), flush=True)
    assert e32 <= 1.1 * e64
