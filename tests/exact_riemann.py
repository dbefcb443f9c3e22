"""Exact solution of the 1-D Euler Riemann problem (ideal gas), following the
textbook construction (Toro, Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., ch. 4): Newton iteration for p*, then sampling at x/t."""
import math

import numpy as np


def _f(p, rhoK, pK, aK, g):
    if p > pK:  # shock
        A = 2.0 / ((g + 1) * rhoK)
        B = (g - 1) / (g + 1) * pK
        f = (p - pK) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1 - (p - pK) / (2 * (B + p)))
    else:  # rarefaction
        f = 2 * aK / (g - 1) * ((p / pK) ** ((g - 1) / (2 * g)) - 1)
        df = 1 / (rhoK * aK) * (p / pK) ** (-(g + 1) / (2 * g))
    return f, df


def star_state(WL, WR, g=1.4):
    rL, uL, pL = WL
    rR, uR, pR = WR
    aL, aR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    p = max(1e-8, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (aL + aR))
    for _ in range(100):
        fL, dL = _f(p, rL, pL, aL, g)
        fR, dR = _f(p, rR, pR, aR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        pn = max(pn, 1e-10)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _f(p, rL, pL, aL, g)
    fR, _ = _f(p, rR, pR, aR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(WL, WR, xi, g=1.4):
    """Primitive (rho, u, p) at similarity coordinates xi = (x - x0)/t."""
    rL, uL, pL = WL
    rR, uR, pR = WR
    aL, aR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    ps, us = star_state(WL, WR, g)
    out = np.zeros((len(xi), 3))
    for n, s in enumerate(xi):
        if s <= us:
            if ps > pL:
                SL = uL - aL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if s <= SL:
                    w = (rL, uL, pL)
                else:
                    r = rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    w = (r, us, ps)
            else:
                SHL = uL - aL
                aS = aL * (ps / pL) ** ((g - 1) / (2 * g))
                STL = us - aS
                if s <= SHL:
                    w = (rL, uL, pL)
                elif s >= STL:
                    w = (rL * (ps / pL) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * aL) * (uL - s)
                    w = (rL * c ** (2 / (g - 1)), 2 / (g + 1) * (aL + (g - 1) / 2 * uL + s), pL * c ** (2 * g / (g - 1)))
        else:
            if ps > pR:
                SR = uR + aR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if s >= SR:
                    w = (rR, uR, pR)
                else:
                    r = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    w = (r, us, ps)
            else:
                SHR = uR + aR
                aS = aR * (ps / pR) ** ((g - 1) / (2 * g))
                STR = us + aS
                if s >= SHR:
                    w = (rR, uR, pR)
                elif s <= STR:
                    w = (rR * (ps / pR) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * aR) * (uR - s)
                    w = (rR * c ** (2 / (g - 1)), 2 / (g + 1) * (-aR + (g - 1) / 2 * uR + s), pR * c ** (2 * g / (g - 1)))
        out[n] = w
    return out
