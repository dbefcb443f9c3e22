#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference's own
translation units (oracle/_ref, built by oracle/build_ref.py from the sources
under /root/reference). Run in the development container only:

    python3 tests/golden/make_golden.py

Fixtures are data only (inputs and expected outputs): parsed .inp maps,
limiter values, and small whole-run results (final U_OLD, cell types, scheme
index, dt/time logs) for each slot combination the reference builds with.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

INP = ["sod.inp", "inclinesod.inp", "vortex.inp", "pnggeo.inp", "edge.inp"]


def run_cases():
    cases = []
    for scheme in ["MUSCL", "MHM", "Godunov", "Lax"]:
        for rs in ["HLL", "PVRS", "StegerWarming"]:
            for lim in ["minmod", "VanLeer"]:
                c = O.Case(N_x=32, N_y=12, hi_y=0.25, Scheme=scheme, Riemann=rs, Limiter=lim, case="sodx")
                if O.ref_available(c):
                    cases.append(("sodx_%s_%s_%s" % (scheme, rs, lim), c, 12))
    cases.append(("sodx_MUSCL_HLLC_minmod", O.Case(N_x=32, N_y=12, hi_y=0.25, Riemann="HLLC", case="sodx"), 12))
    cases.append(("sod_rows", O.Case(N_x=24, N_y=24, hi_y=1.0, case="sod"), 12))
    for b in ["IDWBaseGFM", "LinearBaseGFM", "RiemannBaseGFM", "BCGFM", "HBCGFM"]:
        cases.append(("incline_%s" % b, O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary=b, case="incline", Psy_time=0.3),
                      10))
    cases.append(("incline_MHM_IDW", O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM", Scheme="MHM",
                                            case="incline"), 10))
    cases.append(("incline_SW_IDW", O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM",
                                           Riemann="StegerWarming", case="incline"), 10))
    cases.append(("vortex_IDW", O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM", case="vortex",
                                       extra=dict(fa=1.0, f0=0.0, f1=1.0)), 10))
    # PVRS in 2-D: the inclined channel, and a collision along y whose faces
    # take the y solver's branch with the stray store (PVRS/Riemann_solver.cpp:550)
    for s in ["MUSCL", "MHM"]:
        cases.append(("incline_PVRS_%s" % s, O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM", Scheme=s,
                                                    Riemann="PVRS", case="incline"), 10))
        cases.append(("collision_PVRS_%s" % s, pvrs_collision_case(s, 24), 8))
    # SimpleBound (its sources keep the single-device signatures; oracle/ref_simplebound.cpp adapts them)
    cases.append(("incline_SimpleBound", O.Case(N_x=64, N_y=64, hi_y=1.0, Boundary="SimpleBound", case="incline",
                                                Psy_time=0.3), 10))
    cases.append(("vortex_SimpleBound", O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="SimpleBound", case="vortex",
                                               extra=dict(fa=1.0, f0=0.0, f1=1.0)), 10))
    # Lax in 2-D: the row-split Sod hook and the inclined channel with IDW ghost fluid
    cases.append(("sod_Lax", O.Case(N_x=24, N_y=24, hi_y=1.0, Scheme="Lax", case="sod"), 12))
    cases.append(("incline_Lax_IDW", O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM", Scheme="Lax",
                                            case="incline", Psy_time=0.3), 10))
    for s, r, l in [("Godunov", "HLL", "minmod"), ("Godunov", "StegerWarming", "minmod"),
                    ("MUSCL", "HLL", "VanLeer"), ("MHM", "StegerWarming", "minmod"), ("MHM", "HLL", "VanLeer")]:
        cases.append(("incline_%s_%s_%s_IDW" % (s, r, l),
                      O.Case(N_x=28, N_y=28, hi_y=1.0, Boundary="IDWBaseGFM", Scheme=s, Riemann=r, Limiter=l,
                             case="incline", Psy_time=0.3), 10))
    # Test/PNGGeo's shock initialiser and edge hook (adapted signatures) over InclineSod's channel
    cases.append(("pngshock_SW_IDW", O.Case(N_x=40, N_y=40, hi_y=1.0, Boundary="IDWBaseGFM", Riemann="StegerWarming",
                                            case="pngshock", edge="pnggeo", Psy_time=0.3, CFL_number=0.1,
                                            extra=dict(rho_L=1.0, u_L=0.0, v_L=0.0, p_L=0.71429)), 12))
    return cases


def pvrs_collision_case(scheme, n):
    """Row-split hook (Test/Sod/Initialize.cpp): the lower half at rest, the
    upper half moving down supersonically onto it."""
    return O.Case(N_x=n, N_y=n, hi_y=1.0, Scheme=scheme, Riemann="PVRS", case="sod", CFL_number=0.3,
                  extra=dict(rho_L=1.0, u_L=0.3, v_L=0.0, p_L=1.0, rho_R=1.0, u_R=0.0, v_R=-2.0, p_R=2.0))


def case_meta(c):
    return dict(N_x=c.N_x, N_y=c.N_y, hi_x=c.hi_x, hi_y=c.hi_y, Psy_time=c.Psy_time, CFL_number=c.CFL_number,
                Scheme=c.Scheme, Riemann=c.Riemann, Limiter=c.Limiter, Boundary=c.Boundary, case=c.case,
                edge=c.edge, extra=c.extra)


def main():
    if not O.ref_available():
        print("make_golden: oracle/_ref is not built (needs /root/reference); nothing written")
        return 1
    out = {}
    for name in INP:
        p = os.path.join(HERE, "inp", name)
        out[name] = {"numbers": O.parse_ref(p, False), "strings": O.parse_ref(p, True)}
    with open(os.path.join(HERE, "inp_parsed.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)

    rng = np.random.default_rng(7)
    # limiter inputs: random, equal, sign-changing, zero and tie cases
    v = rng.normal(size=(4000, 4))
    v[::7, 1] = v[::7, 0]
    v[::11, 2] = v[::11, 1]
    v[::13] = np.round(v[::13], 1)
    v[::17, :] = 0.0
    v[5::19, 3] = 2 * v[5::19, 2] - v[5::19, 1]
    lim = {"inputs": v}
    for name in ["minmod", "VanLeer"]:
        L = O.ref_lib("libref_sodx_MUSCL_HLL_%s_Regular.so" % name)
        lim[name + "_L"] = np.array([L.ref_lim_L(a, b, c) for a, b, c, _ in v])
        lim[name + "_R"] = np.array([L.ref_lim_R(b, c, d) for _, b, c, d in v])
    np.savez_compressed(os.path.join(HERE, "limiters.npz"), **lim)

    meta = {}
    arrays = {}
    for name, c, n in run_cases():
        r = O.run_ref(c, n)
        meta[name] = dict(case_meta(c), steps=n, steps_taken=r["steps"], dt0=r["dt0"])
        arrays[name + "__U"] = r["U_OLD"]
        arrays[name + "__type"] = r["XY"][5::6].astype(np.uint8)
        arrays[name + "__dt"] = r["dt"]
        arrays[name + "__t"] = r["t"]
        if c.Boundary != "Regular":
            arrays[name + "__S"] = r["S"]
            arrays[name + "__geom"] = r["XY"].reshape(-1, 6)[:, 2:5].copy()
        print("golden", name, r["steps"], flush=True)
    np.savez_compressed(os.path.join(HERE, "runs.npz"), **arrays)
    with open(os.path.join(HERE, "runs.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)

    # WriteData (Src/Utils/WriteData.cpp): the files the reference writes for one
    # state, at iterations 0 and 10 (result.visit gets both entries)
    c, U, XY, t = vtk_case()
    vdir = os.path.join(HERE, "vtk")
    os.makedirs(vdir, exist_ok=True)
    for f in os.listdir(vdir):
        os.remove(os.path.join(vdir, f))
    O.ref_write_vtk(c, U, XY, 0, 0.0, vdir)
    O.ref_write_vtk(c, U, XY, 10, t, vdir)
    np.savez_compressed(os.path.join(HERE, "vtk_state.npz"), U=U, XY=XY, t=np.array([t]))
    print("golden vtk", sorted(os.listdir(vdir)), flush=True)
    return 0


def vtk_case():
    """An inclined-tube state after 10 steps (28 x 20 cells) with a near-vacuum
    density, a non-finite value and a negative pressure planted in fluid cells,
    so the writer's clamp and special-value formatting are covered."""
    c = O.Case(N_x=28, N_y=20, hi_y=1.0, Boundary="IDWBaseGFM", case="incline", Psy_time=0.3)
    r = O.run_ref(c, 10)
    U, XY = r["U_OLD"].copy(), r["XY"].copy()
    fluid = np.flatnonzero(XY[5::6] == 0)
    U[fluid[3] * 4] = 1e-9
    U[fluid[10] * 4 + 1] = np.nan
    U[fluid[20] * 4 + 3] = -5.0
    U[fluid[30] * 4 + 2] = 1.25e7
    return c, U, XY, float(r["t"][r["steps"] - 1])


if __name__ == "__main__":
    sys.exit(main())
