"""Shared helpers for the parity tests: the same problem on the GPU library and
on the CPU oracle, and the comparison rules.

Bars (stated once, used everywhere):
  * cell types and scheme-index records: bit-exact;
  * conserved fields, FP64 path vs the oracle with sqrt sound speeds: bit-exact
    (NaN payloads ignored: every NaN compares equal to every NaN);
  * conserved fields vs the oracle with the reference's pow(x, 0.5): relative
    L-inf <= 1e-12 per conserved variable, relative to that variable's max |.|.
"""
import numpy as np

REL_TOL = 1e-12


def gpu_config(case, **over):
    import ibtfo_amd as ib
    kw = dict(N_x=case.N_x, N_y=case.N_y, lo_x=case.lo_x, lo_y=case.lo_y, hi_x=case.hi_x, hi_y=case.hi_y,
              Psy_time=case.Psy_time, max_iter=case.max_iter, CFL_number=case.CFL_number,
              num_ghost_cell=case.num_ghost_cell, gamma=case.gamma, Scheme=case.Scheme, Riemann=case.Riemann,
              Limiter=case.Limiter, Boundary=case.Boundary, Time=case.Time, Edge=case.edge, quirks=case.quirks,
              edge_state=case.edge_prm())
    kw.update(over)
    return ib.config(**kw)


def gpu_hooks(case):
    """(levelset, init) arguments for ibtfo_amd.run_case matching the oracle's hooks."""
    e = case.extra
    if case.phi is not None:
        ls = ("field", ())
    elif case.case in ("incline", "pngshock"):
        ls = ("channel", case.ls_prm())
    elif case.case == "vortex":
        ls = ("annulus", case.ls_prm())
    else:
        ls = ("none", ())
    init = {"sod": "sod_rows", "sodx": "sod_x", "incline": "incline", "vortex": "vortex", "uniform": "uniform",
            "shu": "shu", "pngshock": "shock"}[case.case]
    return ls, (init, case.init_prm())


def gpu_run(case, nsteps, graph_chunk=0, chunk=64):
    import ibtfo_amd as ib
    cfg = gpu_config(case)
    ls, init = gpu_hooks(case)
    s = ib.Solver(cfg)
    if case.phi is not None:
        s.geometry("field", (), phi=case.phi)
    else:
        s.geometry(ls[0], ls[1])
    s.initialize(init[0], init[1])
    s.prepare(case.init_boundary())
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        s.step(n, graph_chunk)
        it, _, _, _ = s.status()
        if it - done < n:
            done = it
            break
        done = it
    it, t, dt, rc = s.status()
    out = dict(U_OLD=s.state(), XY=s.xycoord(), S=s.scheme_index(), steps=it, t=t, dt=dt, rc=rc, solver=s)
    d, tl, m = s.log(min(it, 4000))
    out.update(dt_log=d, t_log=tl, mass_log=m)
    return out


def nan_equal_bits(a, b):
    """Number of positions whose bits differ, counting any NaN equal to any NaN."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    diff = (a.view(np.uint64) != b.view(np.uint64)) & ~both_nan
    return int(diff.sum())


def rel_linf(a, b, ncomp=4):
    """Per-component max|a-b| / max|b| over finite entries (NaN positions must agree)."""
    a = np.asarray(a).reshape(-1, ncomp)
    b = np.asarray(b).reshape(-1, ncomp)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN pattern differs"
    out = []
    for k in range(ncomp):
        m = np.isfinite(b[:, k])
        scale = np.max(np.abs(b[m, k])) if m.any() else 1.0
        out.append(float(np.max(np.abs(a[m, k] - b[m, k])) / (scale if scale > 0 else 1.0)) if m.any() else 0.0)
    return out


def exact_mass(U, types, dx, dy, alpha=None):
    """CaculateMass of a state summed exactly (math.fsum of the per-cell terms
    rho*dx*dy of fluid cells, alpha*rho*dx*dy of cut cells, each term rounded as
    the kernels round it), to check a reduction's own summation error."""
    import math
    rho = np.asarray(U)[0::4]
    t = np.asarray(types)
    terms = list((rho[t == 0] * dx) * dy)
    if alpha is not None:
        a = np.asarray(alpha)
        terms += list(((a[t == 3] * rho[t == 3]) * dx) * dy)
    return math.fsum(terms)
