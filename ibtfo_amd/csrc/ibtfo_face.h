// Device helpers shared by the step kernels' translation units (the marching
// strip in ibtfo_step.hip, the row-parallel strip in ibtfo_rowpar.hip): wave
// reductions, the dt / mass terms and floors of an updated cell, the block
// partial record, and the face machinery (primitive views, MUSCL-Hancock
// predictor and positivity fallback, the Riemann solvers' dispatch, the x and y
// faces, the limited cell states).
#pragma once
#include "ibtfo_internal.h"

namespace ibtfo {

__device__ __forceinline__ double ld_agent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = (v < w) ? w : v;
    }
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ComputeDt.cpp:34-46 contribution of one fluid cell (std::max keeps the old
// maximum when the candidate is NaN); amax accumulates the squared sound speed.
__device__ __forceinline__ void dt_terms(double gam, double gm1, const double* U, double& amax, double& umax,
                                         double& vmax, int& bad) {
    PState<double> s = pstate(gm1, U);
    double rho = s.rho, u = s.u, v = s.v, p = s.p;
    // the maximum sound speed is taken as sqrt of the maximum a^2: RN(sqrt) is
    // monotone, so this equals the maximum of the rounded sound speeds
    double a2 = dq(gam * p, s.ir);
    amax = (amax < a2) ? a2 : amax;
    double au = ib_fabs(u), av = ib_fabs(v);
    umax = (umax < au) ? au : umax;
    vmax = (vmax < av) ? av : vmax;
    bad |= !(rho > 0.0 && p > 0.0);
}

// The corrected mode's density / pressure floors on an updated fluid cell fused
// with its dt_terms: both need 1/rho, u, v and p, which are evaluated once, in
// pstate's operation order.
__device__ __forceinline__ void floor_and_dt(double gam, double gm1, const Rcp<double>& r_gm1, double* q,
                                             int quirks, double& amax, double& umax, double& vmax, int& bad) {
    if (!quirks && !(q[0] > RHO_FLOOR)) q[0] = RHO_FLOOR;
    const double rho = q[0];
    Rcp<double> ir = rcp_of(rho);
    double u = dq(q[1], ir), v = dq(q[2], ir);
    double ke = 0.5 * rho * (u * u + v * v);
    double p = gm1 * (q[3] - ke);
    if (!quirks && !(p > P_FLOOR)) {
        q[3] = dq(P_FLOOR, r_gm1) + ke;
        p = gm1 * (q[3] - ke);
    }
    double a2 = dq(gam * p, ir);
    amax = (amax < a2) ? a2 : amax;
    double au = ib_fabs(u), av = ib_fabs(v);
    umax = (umax < au) ? au : umax;
    vmax = (vmax < av) ? av : vmax;
    bad |= !(rho > 0.0 && p > 0.0);
}

// Per-block partial of the dt / mass reductions: {max a^2, max |u|, max |v|,
// mass, non-physical flag} at part[PART_STRIDE*bid ..], written with plain
// stores. The one-block k_finish launched after the step reduces them in block
// order (deterministic), so no grid-wide fence or atomic is needed: the kernel
// boundary publishes the partials.
__device__ __forceinline__ void block_partials(double* part, int bid, double amax, double umax, double vmax,
                                               double msum, int bad) {
    amax = wave_max(amax);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    msum = wave_sum(msum);
    bad = __any(bad);
    if (threadIdx.x == 0) {
        double* q = part + PART_STRIDE * (long long)bid;
        q[0] = amax; q[1] = umax; q[2] = vmax; q[3] = msum; q[4] = bad ? 1.0 : 0.0;
    }
}

// Periodic ring written by the step kernel (StepArgs::edge_fused = 2): ghost
// (i, j) holds interior (i +- nx, j +- ny), each coordinate wrapped on its own,
// as the edge kernel copies it (nx, ny >= ng, so every source is interior). The
// store of interior cell (c, j) writes its images across the x edge, the y edge
// and the corner: none for a cell more than ng cells from both edges.
__device__ __forceinline__ void periodic_images(double* __restrict__ Uo, const Layout& L, const int c, const int j,
                                                const double (&e)[4]) {
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    const int xs[3] = {c, c < 2 * ng ? c + nx : -1, c >= nx ? c - nx : -1};
    const int ys[3] = {j, j < 2 * ng ? j + ny : -1, j >= ny ? j - ny : -1};
    if (xs[1] < 0 && xs[2] < 0 && ys[1] < 0 && ys[2] < 0) return;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if ((p == 0 && q == 0) || xs[p] < 0 || ys[q] < 0) continue;
            const long long o = L.at(xs[p], ys[q]);
#pragma unroll
            for (int k = 0; k < 4; k++) Uo[o + k * L.plane] = e[k];
        }
}

// Primitive view of an FP64 conserved state in the compute type T. FP64:
// pstate itself (correctly rounded, the reference's order). Mixed precision:
// 1/rho, u, v and p = (gamma-1)(E - rho |u|^2 / 2) are recovered in FP64 (no
// correct rounding needed) and then rounded to FP32, because near vacuum the
// FP32 difference E - rho |u|^2 / 2 would lose the pressure to cancellation.
template <typename T> __device__ __forceinline__ PState<T> prim_of(double gm1, const double* U);
template <> __device__ __forceinline__ PState<double> prim_of<double>(double gm1, const double* U) {
    return pstate(gm1, U);
}
template <> __device__ __forceinline__ PState<float> prim_of<float>(double gm1, const double* U) {
    const double rho = U[0];
    double y = __builtin_amdgcn_rcp(rho);
    y = __builtin_fma(y, __builtin_fma(-rho, y, 1.0), y);
    const double u = U[1] * y, v = U[2] * y;
    const double p = gm1 * (U[3] - 0.5 * rho * (u * u + v * v));
    PState<float> s;
    s.rho = (float)rho; s.u = (float)u; s.v = (float)v; s.p = (float)p; s.E = (float)U[3];
    s.ir = Rcp<float>{s.rho, (float)y};
    return s;
}

// Face states of one face from the reconstructed FP64 conserved states Ulo/Uhi
// (in/out; the cell averages for Godunov) and the two cell averages u1, u2
// beside it: the MUSCL-Hancock predictor and the corrected mode's positivity
// fallback. T is the compute type of the primitive states and fluxes (double,
// or float in the mixed-precision form).
template <int SCH, bool Y, typename T>
__device__ __forceinline__ void face_states(const GasT<T>& g, double gm1, T dt, int quirks, const double* u1,
                                            const double* u2, double* Ulo, double* Uhi, PState<T>& L,
                                            PState<T>& R) {
    L = prim_of<T>(gm1, Ulo);
    R = prim_of<T>(gm1, Uhi);
    if (SCH == SCH_MHM) {  // MHM_Scheme.cpp:61-96 / :148-183 (same increment on both sides of the face)
        T Fl[4], Fh[4];
        phys_flux<Y>(L, Fl);
        phys_flux<Y>(R, Fh);
        const Rcp<T>& rh = Y ? g.r_dy : g.r_dx;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            T inc = dq(T(0.5) * (Fl[k] - Fh[k]) * dt, rh);
            Ulo[k] = Ulo[k] + inc;
            Uhi[k] = Uhi[k] + inc;
        }
        L = prim_of<T>(gm1, Ulo);
        R = prim_of<T>(gm1, Uhi);
    }
    if (SCH != SCH_GODUNOV && !quirks && !(physical(L) && physical(R))) {
#pragma unroll
        for (int k = 0; k < 4; k++) { Ulo[k] = u1[k]; Uhi[k] = u2[k]; }
        L = prim_of<T>(gm1, Ulo);
        R = prim_of<T>(gm1, Uhi);
    }
}

template <bool Y, int RS, typename T>
__device__ __forceinline__ bool riemann(const GasT<T>& g, const PState<T>& L, const PState<T>& R,
                                        const double* Ulo, const double* Uhi, T* F, bool quirks, bool& stray) {
    T ul[4], uh[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { ul[k] = (T)Ulo[k]; uh[k] = (T)Uhi[k]; }
    stray = false;
    if (RS == RS_SW) { sw_face<Y>(g, L, R, F); return false; }
    if (RS == RS_PVRS) { pvrs_face<Y>(g, ul, uh, F, stray); return false; }
    return hll_face<Y>(g, L, R, ul, uh, F, quirks);
}

// x face between cells u1 (left) and u2 (right), reconstructed states UL, UR
template <int SCH, int RS, typename T>
__device__ __forceinline__ void xface(const GasT<T>& g, double gm1, T dt, int quirks, const double* u1,
                                      const double* u2, double* UL, double* UR, T* F) {
    PState<T> L, R;
    face_states<SCH, false>(g, gm1, dt, quirks, u1, u2, UL, UR, L, R);
    bool stray;
    riemann<false, RS>(g, L, R, UL, UR, F, quirks != 0, stray);
}

// y face between cells u1 (below) and u2 (above); Fx is the lower cell's x-face
// flux in the same row, which is what the reference leaves in the aliased G_OLD
// slot when its y solver stores nothing (HLL/Riemann_solver.cpp:224-226,
// PVRS/Riemann_solver.cpp:550).
template <int SCH, int RS, typename T>
__device__ __forceinline__ void yface(const GasT<T>& g, double gm1, T dt, int quirks, const double* u1,
                                      const double* u2, double* UD, double* UU, const T* Fx, T* G) {
    PState<T> D, Uu;
    face_states<SCH, true>(g, gm1, dt, quirks, u1, u2, UD, UU, D, Uu);
    bool stray;
    bool up = riemann<true, RS>(g, D, Uu, UD, UU, G, quirks != 0, stray);
    if (quirks) {
        if (RS == RS_PVRS && stray) G[3] = Fx[3];
        if (RS == RS_HLL && up) {
#pragma unroll
            for (int k = 0; k < 4; k++) G[k] = Fx[k];
        }
    }
}

// Both reconstructed states of a cell in one direction (cell averages for Godunov)
template <int SCH, int LIM, typename T>
__device__ __forceinline__ void cell_states(const double* um, const double* u0, const double* up, double* lo,
                                            double* hi, int quirks) {
    if (SCH == SCH_GODUNOV) {
#pragma unroll
        for (int k = 0; k < 4; k++) { lo[k] = u0[k]; hi[k] = u0[k]; }
    } else if (sizeof(T) == sizeof(double)) {
        slope_pair<LIM>(um, u0, up, lo, hi, quirks != 0);
    } else {
        // mixed precision: neighbour differences in FP64 (exact cancellation of
        // the common part), the limiter in FP32, the face states back in FP64;
        // corrected-mode limiters only (the form has no reference counterpart)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float a = (float)(u0[k] - um[k]), b = (float)(up[k] - u0[k]);
            float sl, sh;
            if (LIM == LIM_VANLEER) {
                vanleer_pair(a, b, sl, sh);
            } else if (LIM == LIM_MINMOD) {
                sl = lim_minmod(a, b);
                sh = lim_minmod(b, a);
            } else {
                sl = lim_vanalbada(a, b);
                sh = lim_vanalbada(b, a);
            }
            lo[k] = u0[k] + (double)(0.5f * sl);
            hi[k] = u0[k] - (double)(0.5f * sh);
        }
    }
}

}  // namespace ibtfo
