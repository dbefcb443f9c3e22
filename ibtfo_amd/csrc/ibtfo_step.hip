// Fused per-timestep kernels of the IBTFO Euler update for gfx950.
//
// k_step: one launch = TimeAdvance (reconstruction + Riemann flux + conservative
// update, x then y) + the reference Global_Boundary ring + the ComputeDt and
// CaculateMass reductions of the new state + the host loop's bookkeeping
// (now_t += dt, iter++, stop test), for MUSCL / MUSCL-Hancock / Godunov with
// HLL(=HLLC) / PVRS / Steger-Warming and minmod / VanLeer / VanAlbada.
//
// Work decomposition: a 64-lane block owns a strip of 60 interior columns and
// `rows` interior rows and marches upward through it. Lane t holds column
// c = cb - 2 + t (lanes 0, 1, 62 and 63 are the strip's halo columns). Per
// row, each lane reconstructs both face states of its own cell in x
// (neighbours through LDS) and in y (a 3-row register window), so each limiter
// pair is evaluated once; it takes the right state of face c+1/2 from lane t+1
// and the x flux of face c-1/2 from lane t-1 through LDS, and carries the y
// flux of face j-1/2 and its cell's upper y state from the previous row in its
// own LDS slots. A strip reads its cells plus two halo rows above and below
// and two halo columns on each side, and writes each of its cells once.
#include <type_traits>

#include "ibtfo_internal.h"
#include "ibtfo_face.h"

namespace ibtfo {

// ComputeDt (ComputeDt.cpp:23-49) and CaculateMass (CaculateMass.cpp:27-38) from
// the block partials, then the host loop's bookkeeping of Test/Sod/Sod.cpp:154-177:
// dt for the next step, now_t += dt, iter++, stop test, log entry. `advance` = 0
// for the stand-alone ComputeDt (no step taken; `reset` starts the run clock).
// record q of the partials (part, then part2)
__device__ __forceinline__ const double* part_rec(const FinishArgs& f, int q) {
    return q < f.nblocks ? f.part + PART_STRIDE * (long long)q : f.part2 + PART_STRIDE * (long long)(q - f.nblocks);
}

// ComputeDt's dt from the reduced maxima (A2: max squared sound speed), the
// status flag, and the clock, log and stop test of a step (advance = 1) or the
// stand-alone ComputeDt's mass and clock start (advance = 0). now_t and iter
// are the control block's values before this finish.
__device__ __forceinline__ void finish_clock(const FinishArgs& f, Ctl* ctl, double A2, double Um, double Vm,
                                             double m, int bad, int advance, double now_t, int iter) {
    const double A = ib_sqrt(A2);
    if (bad) ctl->status |= 1;
    double dt = f.cfl * f.dx / (A + ((Um < Vm) ? Vm : Um));
    ctl->dt = dt;
    if (advance) {
        double tn = now_t + dt;
        int it = iter + 1;
        int slot = iter % LOGCAP;
        ctl->log_dt[slot] = dt;
        ctl->log_t[slot] = tn;
        ctl->log_mass[slot] = m;
        ctl->mass_now = m;
        ctl->now_t = tn;
        ctl->iter = it;
        ctl->done = !(tn < f.t_end && it < f.max_iter);
        if (!f.keep_cur) ctl->cur ^= 1;
        ctl->ran = 1;
    } else {
        ctl->mass_now = m;
        if (f.reset) {
            ctl->now_t = 0.0;
            ctl->iter = 0;
            ctl->done = !(0.0 < f.t_end && 0 < f.max_iter);
            ctl->ran = 0;
        }
    }
}

template <int NT> __device__ __forceinline__ void finish_body(const FinishArgs& f, const int advance) {
    Ctl* ctl = f.ctl;
    if (advance && ctl->done) return;  // the step kernel did not run
    if (advance == 2) {
        if (threadIdx.x == 0) {
            ctl->cur ^= 1;
            ctl->ran = 1;
        }
        return;
    }
    __shared__ double sA[NT], sU[NT], sV[NT], sM[NT];
    __shared__ int sB[NT];
    const int t = threadIdx.x;
    double A = 0.0, Um = 0.0, Vm = 0.0, m = 0.0;
    int bad = 0;
    // batches of FB records per thread, all loads issued before any is used, so
    // one memory round trip covers FB records (the loop is latency bound)
    constexpr int FB = 8;
    const int nrec = f.nblocks + (f.part2 ? f.nblocks2 : 0);
    for (int q0 = t; q0 < nrec; q0 += FB * NT) {
        double ra[FB], ru[FB], rv[FB], rm[FB], rb[FB];
#pragma unroll
        for (int b = 0; b < FB; b++) {
            const int q = q0 + b * NT;
            const double* r = part_rec(f, q < nrec ? q : q0);
            ra[b] = r[0]; ru[b] = r[1]; rv[b] = r[2]; rm[b] = r[3]; rb[b] = r[4];
        }
#pragma unroll
        for (int b = 0; b < FB; b++) {
            if (q0 + b * NT >= nrec) break;
            A = (A < ra[b]) ? ra[b] : A;
            Um = (Um < ru[b]) ? ru[b] : Um;
            Vm = (Vm < rv[b]) ? rv[b] : Vm;
            m += rm[b];
            bad |= rb[b] != 0.0;
        }
    }
    sA[t] = A; sU[t] = Um; sV[t] = Vm; sM[t] = m; sB[t] = bad;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w) {
            sA[t] = (sA[t] < sA[t + w]) ? sA[t + w] : sA[t];
            sU[t] = (sU[t] < sU[t + w]) ? sU[t + w] : sU[t];
            sV[t] = (sV[t] < sV[t + w]) ? sV[t + w] : sV[t];
            sM[t] = sM[t] + sM[t + w];
            sB[t] |= sB[t + w];
        }
        __syncthreads();
    }
    if (t == 0) finish_clock(f, ctl, sA[0], sU[0], sV[0], sM[0], sB[0], advance, ctl->now_t, ctl->iter);
}
__global__ __launch_bounds__(FINISH_THREADS) void k_finish(FinishArgs f) {
    finish_body<FINISH_THREADS>(f, f.advance);
}

// The same finish in one wave, for at most FINISH_SMALL_MAX records (small
// grids, whose step is a chain of latencies): each lane reads its records in
// one round trip, the control fields are read beside them rather than first,
// and the lanes reduce through cross-lane shuffles instead of an LDS tree.
constexpr int FINISH_SMALL_PER_LANE = FINISH_SMALL_MAX / 64;
__global__ __launch_bounds__(64) void k_finish_small(FinishArgs f) {
    Ctl* ctl = f.ctl;
    const int advance = f.advance;
    if (advance == 2) {
        if (threadIdx.x == 0) {
            if (!ctl->done) {
                ctl->cur ^= 1;
                ctl->ran = 1;
            }
        }
        return;
    }
    const int t = threadIdx.x;
    const int nrec = f.nblocks + (f.part2 ? f.nblocks2 : 0);
    const int done = ctl->done;
    const double now_t = ctl->now_t;
    const int iter = ctl->iter;
    double ra[FINISH_SMALL_PER_LANE], ru[FINISH_SMALL_PER_LANE], rv[FINISH_SMALL_PER_LANE], rm[FINISH_SMALL_PER_LANE],
        rb[FINISH_SMALL_PER_LANE];
#pragma unroll
    for (int b = 0; b < FINISH_SMALL_PER_LANE; b++) {
        const int q = t + 64 * b;
        const double* r = part_rec(f, q < nrec ? q : 0);
        ra[b] = r[0]; ru[b] = r[1]; rv[b] = r[2]; rm[b] = r[3]; rb[b] = r[4];
    }
    double A = 0.0, Um = 0.0, Vm = 0.0, m = 0.0;
    int bad = 0;
#pragma unroll
    for (int b = 0; b < FINISH_SMALL_PER_LANE; b++) {
        if (t + 64 * b >= nrec) break;
        A = (A < ra[b]) ? ra[b] : A;
        Um = (Um < ru[b]) ? ru[b] : Um;
        Vm = (Vm < rv[b]) ? rv[b] : Vm;
        m += rm[b];
        bad |= rb[b] != 0.0;
    }
    A = wave_max(A);
    Um = wave_max(Um);
    Vm = wave_max(Vm);
    m = wave_sum(m);
    bad = __any(bad);
    if (advance && done) return;  // the step kernel did not run
    if (t == 0) finish_clock(f, ctl, A, Um, Vm, m, bad, advance, now_t, iter);
}

// Partials of a 256-thread block: wave reductions, then the four waves in order.
__device__ __forceinline__ void block4_partials(double* part, int bid, double amax, double umax, double vmax,
                                                double msum, int bad) {
    __shared__ double sw[4][4];
    __shared__ int sb[4];
    amax = wave_max(amax);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    msum = wave_sum(msum);
    bad = __any(bad);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sw[w][0] = amax; sw[w][1] = umax; sw[w][2] = vmax; sw[w][3] = msum; sb[w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = sw[0][0], U = sw[0][1], V = sw[0][2], m = sw[0][3];
        int b = sb[0];
        for (int q = 1; q < 4; q++) {
            A = (A < sw[q][0]) ? sw[q][0] : A;
            U = (U < sw[q][1]) ? sw[q][1] : U;
            V = (V < sw[q][2]) ? sw[q][2] : V;
            m = m + sw[q][3];
            b |= sb[q];
        }
        double* r = part + PART_STRIDE * (long long)bid;
        r[0] = A; r[1] = U; r[2] = V; r[3] = m; r[4] = b ? 1.0 : 0.0;
    }
}

// Two-level finish in one launch: block g folds partials [g * fold_per,
// (g + 1) * fold_per) into fold record g; the last block to finish (counted in
// Ctl::folded, with device-scope fences around the count) reduces the fold
// records and runs the clock. Only the records cross blocks, so the fences
// write back and invalidate little.
__global__ __launch_bounds__(FOLD_THREADS) void k_fold(FinishArgs f) {
    if (f.ctl->done) return;  // the step kernel did not run (every block sees the same flag)
    const int nrec = f.nblocks + (f.part2 ? f.nblocks2 : 0);
    const int q0 = blockIdx.x * f.fold_per, q1 = min(q0 + f.fold_per, nrec);
    double A = 0.0, U = 0.0, V = 0.0, m = 0.0;
    int bad = 0;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += FOLD_THREADS) {
        const double* r = part_rec(f, q);
        A = (A < r[0]) ? r[0] : A;
        U = (U < r[1]) ? r[1] : U;
        V = (V < r[2]) ? r[2] : V;
        m += r[3];
        bad |= r[4] != 0.0;
    }
    block4_partials(f.fold, blockIdx.x, A, U, V, m, bad);
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();  // this block's record before its count
        last = atomicAdd(&f.ctl->folded, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every record after the count
    if (threadIdx.x == 0) f.ctl->folded = 0;  // no other block touches it any more
    FinishArgs g = f;
    g.part = f.fold;
    g.nblocks = gridDim.x;
    g.part2 = nullptr;
    g.nblocks2 = 0;
    finish_body<FOLD_THREADS>(g, 1);
}

// Batched finish: block s finishes solve s (slots with a null control block are idle).
__global__ __launch_bounds__(FINISH_THREADS) void k_finish_batch(const FinishArgs* __restrict__ f, int nslots,
                                                                 int advance) {
    if ((int)blockIdx.x >= nslots || f[blockIdx.x].ctl == nullptr) return;
    finish_body<FINISH_THREADS>(f[blockIdx.x], advance);
}

// The step body for one 64-lane block: strip (bx, by) of the solve described
// by `a`; `stage` as StepArgs::stage. Used by the single-solve kernel and by the
// batched kernel, whose blocks each pick their solve from a block map.
// PER: the periodic ring of the output is written beside the stores (edge_fused = 2;
// its own kernel, so that the other kernels' register allocation is not touched).
template <int SCH, int RS, int LIM, typename T, bool CUT, bool PER = false>
__device__ __forceinline__ void step_strip(const StepArgs& a, const int bx, const int by, const int stage) {
    constexpr bool FP64 = sizeof(T) == sizeof(double);
    Ctl* ctl = a.ctl;
    const int lane = threadIdx.x;
    if (ctl->done) {
        if (bx == 0 && by == 0 && lane == 0) ctl->ran = 0;
        return;
    }
    const Layout L = a.L;
    const int cur = ctl->cur ^ (stage == 3 ? 1 : 0);  // the input buffer
    const double dt = ctl->dt;
    const double* __restrict__ Ui = a.buf[cur];
    double* __restrict__ Uo = a.buf[cur ^ 1];
    const long long PL = L.plane;
    GasT<T> g;
    g.gam = (T)a.gam; g.gm1 = (T)a.gm1; g.gp1 = (T)a.gp1; g.c1 = (T)a.c1;
    g.r_gam = Rcp<T>{(T)a.gam, (T)a.r_gam};
    g.r_gm1 = Rcp<T>{(T)a.gm1, (T)a.r_gm1};
    g.r_dx = Rcp<T>{(T)a.dx, (T)a.r_dx};
    g.r_dy = Rcp<T>{(T)a.dy, (T)a.r_dy};
    const T dtc = (T)dt;
    const Rcp<double> r_dx{a.dx, a.r_dx}, r_dy{a.dy, a.r_dy}, r_gm1{a.gm1, a.r_gm1};
    const double dtdx = dt * a.r_dx, dtdy = dt * a.r_dy;  // mixed form only
    const double dx = a.dx, dy = a.dy;
    const int ng = L.ng, nx = L.nx, ny = L.ny;

    const int cb = ng + STRIP_W * bx;  // first interior column of the strip
    const int c = cb - 2 + lane;
    const int cl = min(max(c, 0), L.M - 1);
    const int j0 = ng + a.rows * by;
    const int j1 = min(j0 + a.rows, ny + ng);

    __shared__ double rb[2][4][64];  // row values of columns cb-2 .. cb+61
    __shared__ double hb[4][64];     // lower-face states of the x pairs
    __shared__ T fb[4][64];          // x-face fluxes
    // carried from one row to the next, per lane: kept in LDS rather than in
    // registers (fewer VGPRs: four waves per SIMD instead of three)
    __shared__ T gpb[4][64];         // y flux of the lane's cell's lower face
    __shared__ double ylb[4][64];    // the lane's cell state on its upper y face

    // reference Global_Boundary corners (Test/Sod/Global_Boudary.cpp:62-75 applied to
    // ring rows): U_OLD(i<ng or i>=N_x+ng, j ring) = U_NEW(ng or N_x+ng-1, j)
    if (a.edge_fused == 1 && bx == 0 && by == 0 && lane < 4 * ng * ng) {
        int side = lane / (2 * ng * ng);
        int rem = lane % (2 * ng * ng);
        int ri = rem / ng, cc = rem % ng;
        int row = ri < ng ? ri : ny + ri;
        int col = side == 0 ? cc : nx + ng + cc;
        int scol = side == 0 ? ng : nx + ng - 1;
        long long o = L.at(col, row);
        for (int k = 0; k < 4; k++) {
            double v = a.corner_from_input ? Ui[L.at(scol, row) + k * PL] : a.corner[(side * 2 * ng + ri) * 4 + k];
            Uo[o + k * PL] = v;
        }
    }

    // column window: w[0..2] = rows j, j+1, j+2 of column c at iteration j;
    // ylb = the y state of row j on its upper face (from the previous pair)
    double w[3][4];
    {
        double wm[4];
        long long o = L.at(cl, j0 - 2);
#pragma unroll
        for (int k = 0; k < 4; k++) wm[k] = Ui[o + k * PL];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            o = L.at(cl, j0 - 1 + r);
#pragma unroll
            for (int k = 0; k < 4; k++) w[r][k] = Ui[o + k * PL];
        }
        double hi_unused[4], ylo[4];
        cell_states<SCH, LIM, T>(wm, w[0], w[1], ylo, hi_unused, a.quirks);
#pragma unroll
        for (int k = 0; k < 4; k++) { ylb[k][lane] = ylo[k]; gpb[k][lane] = T(0); }
    }

    double amax = 0.0, umax = 0.0, vmax = 0.0, msum = 0.0;
    int bad = 0;
    const bool upd_col = lane >= 2 && lane < STRIP_W + 2 && c < nx + ng;
    const int ln = min(lane + 1, 63), lp = max(lane - 1, 0);

    for (int j = j0 - 1; j < j1; j++) {
        const int b = j & 1;
        // prefetch the next window row and next halo value (none after the last
        // row: they would only be discarded)
        double nw[4];
        if (j + 1 < j1) {
            long long o = L.at(cl, min(j + 3, L.N - 1));
#pragma unroll
            for (int k = 0; k < 4; k++) nw[k] = Ui[o + k * PL];
        }
        T F[4] = {0, 0, 0, 0};
        // x faces of row j; the first iteration (j = j0-1) only needs the y face
        // below row j0, unless that face may take the aliased x flux
        if (j >= j0 || (a.quirks && RS != RS_SW)) {
#pragma unroll
            for (int k = 0; k < 4; k++) rb[b][k][lane] = w[0][k];
            __syncthreads();
            double xm[4], xp[4], xlo[4], xhi[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                xm[k] = rb[b][k][lp];
                xp[k] = rb[b][k][ln];
            }
            cell_states<SCH, LIM, T>(xm, w[0], xp, xlo, xhi, a.quirks);
#pragma unroll
            for (int k = 0; k < 4; k++) hb[k][lane] = xhi[k];
            __syncthreads();
            double xr[4];
#pragma unroll
            for (int k = 0; k < 4; k++) xr[k] = hb[k][ln];
            xface<SCH, RS>(g, a.gm1, dtc, a.quirks, w[0], xp, xlo, xr, F);
        }
        // y face j+1/2: this row's upper state and the lower state of row j+1
        double ynlo[4], ynhi[4];
        T G[4];
        cell_states<SCH, LIM, T>(w[0], w[1], w[2], ynlo, ynhi, a.quirks);
        {
            double ylo[4];
#pragma unroll
            for (int k = 0; k < 4; k++) ylo[k] = ylb[k][lane];
            yface<SCH, RS>(g, a.gm1, dtc, a.quirks, w[0], w[1], ylo, ynhi, F, G);
        }

        if (j >= j0) {
#pragma unroll
            for (int k = 0; k < 4; k++) fb[k][lane] = F[k];
            __syncthreads();
            if (upd_col) {
                double un[4];
                const long long o = L.at(c, j);
                const uint8_t tc = a.type[o];
                const bool fluid = CUT ? (tc & 3) == 0 : tc == 0;
                const bool merged = CUT && (tc & MERGED_BIT);
                if (fluid) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (FP64) {
                            double Fm = fb[k][lane - 1];
                            double v = w[0][k] - dq(dt * ((double)F[k] - Fm), r_dx);
                            un[k] = v - dq(dt * ((double)G[k] - (double)gpb[k][lane]), r_dy);
                        } else {  // FP32 fluxes, FP64 accumulation
                            double Fm = (double)fb[k][lane - 1];
                            un[k] = w[0][k] - dtdx * ((double)F[k] - Fm) - dtdy * ((double)G[k] - (double)gpb[k][lane]);
                        }
                    }
                    if (stage == 0) {
                        if (!merged) {
                            floor_and_dt(a.gam, a.gm1, r_gm1, un, a.quirks, amax, umax, vmax, bad);
                        } else {  // a merge target: floored here, its partials come from the merge
                            double d0 = 0.0, d1 = 0.0, d2 = 0.0;
                            int db = 0;
                            floor_and_dt(a.gam, a.gm1, r_gm1, un, a.quirks, d0, d1, d2, db);
                        }
                    } else {
                        // RK2 (Heun): each stage's update takes the floors; the second
                        // stage then averages with U^n, which the output buffer holds
                        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
                        int db = 0;
                        floor_and_dt(a.gam, a.gm1, r_gm1, un, a.quirks, d0, d1, d2, db);
                        if (stage >= 2) {
#pragma unroll
                            for (int k = 0; k < 4; k++) un[k] = 0.5 * (Uo[o + k * PL] + un[k]);
                            dt_terms(a.gam, a.gm1, un, amax, umax, vmax, bad);
                        }
                    }
                    if (!merged) msum += un[0] * dx * dy;
                } else if (CUT && (tc & 3) == 3) {
                    // cut cell: aperture-weighted face fluxes and the wall's pressure
                    // force S p (S closes the cell's faces), over the fluid volume
                    // alpha dx dy; small cells are merged (and floored) afterwards
                    const CutCell cc = a.cut[a.cut_index[o]];
                    const double rho = w[0][0], u0 = w[0][1] / rho, v0 = w[0][2] / rho;
                    const double pw = a.gm1 * (w[0][3] - 0.5 * rho * (u0 * u0 + v0 * v0));
                    const double wall[4] = {0.0, pw * cc.sbx, pw * cc.sby, 0.0};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const double Fk = (double)F[k], Fm = (double)fb[k][lane - 1];
                        const double Gk = (double)G[k], Gm = (double)gpb[k][lane];
                        double v = w[0][k] - dt * ((cc.ae * Fk - cc.aw * Fm) * dy) / (cc.alpha * dx * dy);
                        un[k] = v - dt * ((cc.an * Gk - cc.as * Gm) * dx + wall[k]) / (cc.alpha * dx * dy);
                    }
                    // a merge-group cell (small cell or target) takes its floors and
                    // partials from the merge; the others are floored and counted here
                    if (!merged) {
                        floor_and_dt(a.gam, a.gm1, r_gm1, un, a.quirks, amax, umax, vmax, bad);
                        msum += cc.alpha * un[0] * dx * dy;
                    } else if (cc.alpha >= 0.5) {
                        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
                        int db = 0;
                        floor_and_dt(a.gam, a.gm1, r_gm1, un, a.quirks, d0, d1, d2, db);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++) un[k] = w[0][k];
                }
#pragma unroll
                for (int k = 0; k < 4; k++) Uo[o + k * PL] = un[k];
                if (a.edge_fused == 1) {
                    // Test/Sod/Global_Boudary.cpp:24-59 (rows) and :62-75 (columns), which
                    // read U_NEW: the updated state, or for a non-fluid cell under
                    // SimpleBound the value its U_NEW kept
                    double e[4] = {un[0], un[1], un[2], un[3]};
                    if (a.unew_nf && !fluid)
                        for (int k = 0; k < 4; k++) e[k] = a.unew_nf[o + k * PL];
                    if (j < 2 * ng) Uo[L.at(c, 2 * ng - 1 - j) + 2 * PL] = -e[2];
                    if (j == 2 * ng - 1)
                        for (int r = 0; r < ng; r++) {
                            long long q = L.at(c, r);
                            Uo[q] = e[0]; Uo[q + PL] = e[1]; Uo[q + 3 * PL] = e[3];
                        }
                    if (j >= ny) Uo[L.at(c, 2 * (ny + ng) - 1 - j) + 2 * PL] = -e[2];
                    if (j == ny + ng - 1)
                        for (int r = 0; r < ng; r++) {
                            long long q = L.at(c, ny + ng + r);
                            Uo[q] = e[0]; Uo[q + PL] = e[1]; Uo[q + 3 * PL] = e[3];
                        }
                    if (c == ng)
                        for (int r = 0; r < ng; r++) {
                            long long q = L.at(r, j);
                            for (int k = 0; k < 4; k++) Uo[q + k * PL] = e[k];
                        }
                    if (c == nx + ng - 1)
                        for (int r = 0; r < ng; r++) {
                            long long q = L.at(nx + ng + r, j);
                            for (int k = 0; k < 4; k++) Uo[q + k * PL] = e[k];
                        }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            gpb[k][lane] = G[k];
            ylb[k][lane] = ynlo[k];
            w[0][k] = w[1][k];
            w[1][k] = w[2][k];
            w[2][k] = nw[k];
        }
    }
    if (PER && (bx == 0 || bx == a.nbx - 1 || j0 < 2 * ng || j1 > ny)) {
        // the strip's cells within ng of an edge, re-read from this lane's own
        // stores above (after the row loop, where few registers are live)
        for (int j = j0; j < j1; j++) {
            if (!upd_col || (c >= 2 * ng && c < nx && j >= 2 * ng && j < ny)) continue;
            const long long o = L.at(c, j);
            const double e[4] = {Uo[o], Uo[o + PL], Uo[o + 2 * PL], Uo[o + 3 * PL]};
            periodic_images(Uo, L, c, j, e);
        }
    }
    block_partials(a.mass_part, by * a.nbx + bx, amax, umax, vmax, msum, bad);
}

template <int SCH, int RS, int LIM, typename T, bool CUT>
__global__ __launch_bounds__(64) void k_step(StepArgs a) {
    step_strip<SCH, RS, LIM, T, CUT>(a, blockIdx.x, blockIdx.y, a.stage);
}
template <int SCH, int RS, int LIM, typename T>
__global__ __launch_bounds__(64) void k_step_periodic(StepArgs a) {
    step_strip<SCH, RS, LIM, T, false, true>(a, blockIdx.x, blockIdx.y, a.stage);
}

// Batched step: one launch over the strips of many solves. Block b runs strip
// map[b].y of solve map[b].x (row-major over that solve's nbx x nby strips);
// map[b].x < 0 marks an idle block (the grid is sized for the batch capacity so
// that a captured graph stays valid while solves come and go).
template <int SCH, int RS, int LIM, typename T>
__global__ __launch_bounds__(64) void k_step_batch(const StepArgs* __restrict__ slots, const int2* __restrict__ map,
                                                   int stage) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const StepArgs& a = slots[m.x];
    step_strip<SCH, RS, LIM, T, false>(a, m.y % a.nbx, m.y / a.nbx, stage);
}

// Lax-Friedrichs, Lax_Scheme.cpp:27-62: U_new = average of the 4 neighbours minus
// central flux differences. Same strip decomposition; lane t owns column cb-1+t.
__global__ __launch_bounds__(64) void k_step_lax(StepArgs a) {
    Ctl* ctl = a.ctl;
    const int lane = threadIdx.x;
    if (ctl->done) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && lane == 0) ctl->ran = 0;
        return;
    }
    const Layout L = a.L;
    const int cur = ctl->cur;
    const double dt = ctl->dt;
    const double* __restrict__ Ui = a.buf[cur];
    double* __restrict__ Uo = a.buf[cur ^ 1];
    const long long PL = L.plane;
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    const int cb = ng + STRIP_W * blockIdx.x;
    const int c = cb - 1 + lane;
    const int j0 = ng + a.rows * blockIdx.y;
    const int j1 = min(j0 + a.rows, ny + ng);
    double amax = 0.0, umax = 0.0, vmax = 0.0, msum = 0.0;
    int bad = 0;
    if (lane >= 1 && lane <= STRIP_W && c < nx + ng) {
        for (int j = j0; j < j1; j++) {
            const long long o = L.at(c, j);
            double un[4];
            if (a.type[o] == 0) {
                double U[4][4], Fxm[4], Fxp[4], Gym[4], Gyp[4];
                const long long nb[4] = {L.at(c - 1, j), L.at(c + 1, j), L.at(c, j + 1), L.at(c, j - 1)};
                for (int q = 0; q < 4; q++)
                    for (int k = 0; k < 4; k++) U[q][k] = Ui[nb[q] + k * PL];
                flux_of<false>(a.gm1, U[0], Fxm);
                flux_of<false>(a.gm1, U[1], Fxp);
                flux_of<true>(a.gm1, U[2], Gyp);
                flux_of<true>(a.gm1, U[3], Gym);
                for (int k = 0; k < 4; k++)
                    un[k] = 0.25 * (U[0][k] + U[1][k] + U[2][k] + U[3][k]) - 0.5 * dt * (Fxp[k] - Fxm[k]) / a.dx -
                            0.5 * dt * (Gyp[k] - Gym[k]) / a.dy;
                dt_terms(a.gam, a.gm1, un, amax, umax, vmax, bad);
                msum += un[0] * a.dx * a.dy;
            } else {
                for (int k = 0; k < 4; k++) un[k] = Ui[o + k * PL];
            }
            for (int k = 0; k < 4; k++) Uo[o + k * PL] = un[k];
        }
    }
    // the reference edge ring for Lax runs as the separate edge kernel
    block_partials(a.mass_part, blockIdx.y * a.nbx + blockIdx.x, amax, umax, vmax, msum, bad);
}

// ComputeDt + CaculateMass of the current state (ComputeDt.cpp:23-49,
// CaculateMass.cpp:27-38); with `reset`, also starts the run clock.
__global__ __launch_bounds__(64) void k_reduce(ReduceArgs a) {
    const Layout L = a.L;
    const double* Us = a.cur_ctl ? a.ubuf[a.cur_ctl->cur] : a.U;
    const int nint = L.nx * L.ny;
    double amax = 0.0, umax = 0.0, vmax = 0.0, msum = 0.0;
    int bad = 0;
    for (int q = blockIdx.x * 64 + threadIdx.x; q < nint; q += a.nblocks * 64) {
        int i = L.ng + q % L.nx, j = L.ng + q / L.nx;
        long long o = L.at(i, j);
        const uint8_t t = a.type[o];
        if (t == 0 || (t == 3 && a.cut)) {
            double U[4] = {Us[o], Us[o + L.plane], Us[o + 2 * L.plane], Us[o + 3 * L.plane]};
            dt_terms(a.gam, a.gm1, U, amax, umax, vmax, bad);
            if (t == 0) msum += U[0] * a.dx * a.dy;
            else msum += a.cut[a.cut_index[o]].alpha * U[0] * a.dx * a.dy;
        }
    }
    block_partials(a.mass_part, blockIdx.x, amax, umax, vmax, msum, bad);
}

// Small-cell merging (cut cells): each group's target and members take the
// volume-fraction-weighted mean of their updated states; the corrected-mode
// floors, which the small cells skipped in the step, apply to the mean. The
// groups' cells are left out of the step kernel's dt / mass partials; each
// 64-thread block writes theirs, over the merged state, as one more partial
// record (mass: the group's volume times the merged density).
__global__ __launch_bounds__(MERGE_THREADS) void k_merge(MergeArgs a) {
    if (a.ctl->done) return;  // no step was taken (every block sees the same flag)
    const int gi = blockIdx.x * MERGE_THREADS + threadIdx.x;
    double amax = 0.0, umax = 0.0, vmax = 0.0, msum = 0.0;
    int bad = 0;
    if (gi < a.ngroups) {
        const MergeGroup gr = a.groups[gi];
        double* U = a.buf[a.ctl->cur ^ 1];
        const long long PL = a.plane;
        // every load first (absent members read the target, unused), then the
        // sums in group order
        double ut[4], um[4][4];
#pragma unroll
        for (int k = 0; k < 4; k++) ut[k] = U[gr.target + k * PL];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long long o = q < gr.n ? gr.member[q] : gr.target;
#pragma unroll
            for (int k = 0; k < 4; k++) um[q][k] = U[o + k * PL];
        }
        double asum = gr.at;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q < gr.n) asum = asum + gr.am[q];
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double s = gr.at * ut[k];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (q < gr.n) s = s + gr.am[q] * um[q][k];
            v[k] = s / asum;
        }
        floor_and_dt(a.gm1 + 1.0, a.gm1, Rcp<double>{a.gm1, a.r_gm1}, v, a.quirks, amax, umax, vmax, bad);
        msum = gr.at * v[0] * a.dx * a.dy;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q < gr.n) msum = msum + gr.am[q] * v[0] * a.dx * a.dy;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            U[gr.target + k * PL] = v[k];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (q < gr.n) U[gr.member[q] + k * PL] = v[k];
        }
    }
    block_partials(a.part, blockIdx.x, amax, umax, vmax, msum, bad);
}

// Ghost-cell fill, phase 1: values from the donors of U_NEW (the state right
// after the update, before any ghost is rewritten) into a compact buffer.
// IDWBaseGFM/GFM_Bound.cpp:65-106, LinearBaseGFM/GFM_Bound.cpp:58-575,
// BCGFM/BCGFM_Bound.cpp:81-175, HBCGFM/HBCGFM_Bound.cpp:163-178.
__device__ __forceinline__ void fill_one(const FillArgs& a, const int gi) {
    if (a.require_ran && !a.ctl->ran) return;
    if (gi >= a.nghost) return;
    const Ghost gh = a.ghosts[gi];
    if (a.bc == IBTFO_SIMPLE_BOUND && gh.pick != a.level) return;
    const int cur = a.use_cur ? a.ctl->cur : a.fixed;
    const double* U = a.buf[cur];
    // Boundary runs before Global_Boundary in the reference, so a donor in the
    // domain ring still holds the ring of the previous state
    const double* Uring = a.ring_prev ? a.buf[cur ^ 1] : U;
    const long long PL = a.L.plane, P = a.L.pitch;
    const double gm1 = a.gm1;
    auto prim = [&](long long d, bool ring, double& r, double& u, double& v, double& p) {
        const double* S = ring ? Uring : U;
        double q[4] = {S[d], S[d + PL], S[d + 2 * PL], S[d + 3 * PL]};
        primitives(gm1, q, r, u, v, p);
    };
    auto idw = [&](long long d1, int rbits, const double* w, double& R, double& Uu, double& V, double& Pp) {
        double r[4], u[4], v[4], p[4];
        const long long dd[4] = {d1, d1 + P, d1 + P + 1, d1 + 1};
        for (int q = 0; q < 4; q++) prim(dd[q], (rbits >> q) & 1, r[q], u[q], v[q], p[q]);
        if (!a.quirks)  // corrected mode: a donor at the point itself (1/0 weight) is the value
            for (int q = 0; q < 4; q++)
                if (__builtin_isinf(w[q])) { R = r[q]; Uu = u[q]; V = v[q]; Pp = p[q]; return; }
        double s = w[0] + w[1] + w[2] + w[3];
        R = (w[0] * r[0] + w[1] * r[1] + w[2] * r[2] + w[3] * r[3]) / s;
        Uu = (w[0] * u[0] + w[1] * u[1] + w[2] * u[2] + w[3] * u[3]) / s;
        V = (w[0] * v[0] + w[1] * v[1] + w[2] * v[2] + w[3] * v[3]) / s;
        Pp = (w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] * p[3]) / s;
    };
    const double n_x = gh.nx, n_y = gh.ny;
    double rw, uw, vw, pw;
    if (a.bc == 1) {  // IDWBaseGFM
        if (!a.quirks && gh.a[0] + gh.a[1] + gh.a[2] + gh.a[3] == 0.0) {
            // corrected mode: no fluid donor, the ghost keeps its value
            double* o = a.out + 4 * (long long)gi;
            for (int k = 0; k < 4; k++) o[k] = U[gh.cell + k * PL];
            return;
        }
        idw(gh.d1, gh.ring & 15, gh.a, rw, uw, vw, pw);
        double u_n = uw * n_x + vw * n_y;
        double u_t = uw * n_y - vw * n_x;
        u_n = -u_n;
        uw = u_n * n_x + u_t * n_y;
        vw = u_n * n_y - u_t * n_x;
    } else if (a.bc == IBTFO_SIMPLE_BOUND) {
        // SimpleBound/LS_SimpleBound.cpp:38-105: reflect the x donor's u and/or the y
        // donor's v (the mean of both reflections when both exist). A donor that is
        // a ghost filled earlier in the reference's order is read from its new
        // value (an earlier dependency level), any other from the state.
        const int IDX = gh.idx, IDY = gh.idy;
        if (IDX == 0 && IDY == 0) {  // no fluid within reach: the ghost keeps its value
            double* o = a.out + 4 * (long long)gi;
            for (int k = 0; k < 4; k++) o[k] = U[gh.cell + k * PL];
            return;
        }
        auto donor = [&](long long d, int earlier, bool ring, double& r, double& u, double& v, double& p) {
            if (earlier >= 0) {
                const double* q = a.out + 4 * (long long)earlier;
                double qq[4] = {q[0], q[1], q[2], q[3]};
                primitives(gm1, qq, r, u, v, p);
            } else {
                prim(d, ring, r, u, v, p);
            }
        };
        double rx = 0, ux = 0, vx = 0, px = 0, ry = 0, uy = 0, vy = 0, py = 0;
        if (IDX != 0) donor(gh.d1, gh.idxe, gh.ring & 1, rx, ux, vx, px);
        if (IDY != 0) donor(gh.d1e, gh.idye, (gh.ring >> 4) & 1, ry, uy, vy, py);
        if (IDX != 0 && IDY != 0) {
            rw = 0.5 * (rx + ry); uw = 0.5 * (-ux + uy); vw = 0.5 * (vx - vy); pw = 0.5 * (px + py);
        } else if (IDX != 0) {
            rw = rx; uw = -ux; vw = vx; pw = px;
        } else {
            rw = ry; uw = uy; vw = -vy; pw = py;
        }
    } else if (a.bc == 2 || a.bc == 3) {  // Linear/RiemannBaseGFM: one donor
        int pick = (int)gh.a[0];
        if (pick >= 0) {
            const long long dd[4] = {gh.d1, gh.d1 + P, gh.d1 + P + 1, gh.d1 + 1};
            prim(dd[pick], (gh.ring >> pick) & 1, rw, uw, vw, pw);
        } else {
            rw = 0; uw = 0; vw = 0; pw = 0;
        }
        double u_n = uw * n_x + vw * n_y;
        double u_t = uw * n_y - vw * n_x;
        u_n = -u_n;
        uw = u_n * n_x + u_t * n_y;
        vw = u_n * n_y - u_t * n_x;
    } else {  // BCGFM / HBCGFM
        double delta = ib_sqrt(a.dxg * a.dxg + a.dyg * a.dyg) / 2;
        double R_E = delta, R_G = gh.phi;
        double ri, ui, vi, pi_, re, ue, ve, pe;
        idw(gh.d1, gh.ring & 15, gh.a, ri, ui, vi, pi_);
        double uni = ui * n_x + vi * n_y, uti = ui * n_y - vi * n_x;
        idw(gh.d1e, (gh.ring >> 4) & 15, gh.ae, re, ue, ve, pe);
        double une = ue * n_x + ve * n_y, ute = ue * n_y - ve * n_x;
        double un, ut;
        if (a.bc == 5 && R_G > delta) {
            rw = ri; pw = pi_; ut = uti; un = -uni;
        } else {
            double A = ((delta + R_E) * (delta + R_E) - R_G * R_G), B = (delta * delta - R_G * R_G);
            double D = (R_E * R_E + 2 * delta * R_E);
            rw = (ri * A - re * B) / D;
            pw = (pi_ * A - pe * B) / D;
            ut = (uti * A - ute * B) / D;
            un = -(uni * A - une * B) / D;
        }
        uw = un * n_x + ut * n_y;
        vw = un * n_y - ut * n_x;
    }
    double* o = a.out + 4 * (long long)gi;
    o[0] = rw;
    o[1] = rw * uw;
    o[2] = rw * vw;
    o[3] = pw / gm1 + 0.5 * rw * (uw * uw + vw * vw);
}
__global__ void k_fill(FillArgs a) { fill_one(a, blockIdx.x * blockDim.x + threadIdx.x); }
// Batched fill: block b covers ghosts map[b].y .. +FILL_THREADS-1 of solve map[b].x.
__global__ void k_fill_batch(const FillArgs* __restrict__ slots, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x >= 0) fill_one(slots[m.x], m.y + threadIdx.x);
}

// Ghost-cell fill, phase 2: write the values into the current state.
__device__ __forceinline__ void scatter_one(const FillArgs& a, const int gi) {
    if (a.require_ran && !a.ctl->ran) return;
    if (gi >= a.nghost) return;
    double* U = a.buf[a.use_cur ? a.ctl->cur : a.fixed];
    const long long c = a.ghosts[gi].cell, PL = a.L.plane;
    const double* v = a.out + 4 * (long long)gi;
    U[c] = v[0];
    U[c + PL] = v[1];
    U[c + 2 * PL] = v[2];
    U[c + 3 * PL] = v[3];
}
__global__ void k_scatter(FillArgs a) { scatter_one(a, blockIdx.x * blockDim.x + threadIdx.x); }
__global__ void k_scatter_batch(const FillArgs* __restrict__ slots, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x >= 0) scatter_one(slots[m.x], m.y + threadIdx.x);
}

// Ring of the current state for the edge policies that are not fused into the
// step (the reference policy for Lax; periodic; inflow). Reads the interior of
// the current state, writes its ring; one thread per ring cell.
__device__ __forceinline__ void edge_one(const EdgeArgs& a, const int q) {
    if (a.require_ran && !a.ctl->ran) return;
    const Layout L = a.L;
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    const int nring = L.M * L.N - nx * ny;
    if (q >= nring) return;
    // enumerate ring cells: bottom rows, top rows, then left/right columns of interior rows
    int i, j;
    int nb = L.M * ng;
    if (q < nb) { j = q / L.M; i = q % L.M; }
    else if (q < 2 * nb) { int r = q - nb; j = ny + ng + r / L.M; i = r % L.M; }
    else { int r = q - 2 * nb; j = ng + r / (2 * ng); int s = r % (2 * ng); i = s < ng ? s : nx + s; }
    double* U = a.buf[a.ctl->cur];
    const long long PL = L.plane;
    double v[4];
    if (a.kind == 1) {  // periodic
        int si = i < ng ? i + nx : (i >= nx + ng ? i - nx : i);
        int sj = j < ng ? j + ny : (j >= ny + ng ? j - ny : j);
        long long s = L.at(si, sj);
        const double* S = (a.unew_nf && a.type[s] != 0) ? a.unew_nf : U;
        for (int k = 0; k < 4; k++) v[k] = S[s + k * PL];
    } else if (a.kind == 3) {
        // Test/PNGGeo/Global_Boudary.cpp:14-79 (U_NEW = the state before the ghost
        // fill's phase 2): rows as the Sod hook over the interior columns, the
        // right columns zero-gradient, the left columns held at the inflow state;
        // the ring corners are not written
        const bool row = j < ng || j >= ny + ng;
        if (row && (i < ng || i >= nx + ng)) return;
        auto src = [&](long long s, int k) {
            const double* S = (a.unew_nf && a.type[s] != 0) ? a.unew_nf : U;
            return S[s + k * PL];
        };
        if (row) {
            const int sr = j < ng ? 2 * ng - 1 : ny + ng - 1;
            const int rr = j < ng ? 2 * ng - j - 1 : 2 * (ny + ng) - 1 - j;
            for (int k = 0; k < 4; k++) v[k] = src(L.at(i, sr), k);
            v[2] = -src(L.at(i, rr), 2);
        } else if (i < ng) {
            for (int k = 0; k < 4; k++) v[k] = a.inflow[k];
        } else {
            for (int k = 0; k < 4; k++) v[k] = src(L.at(nx + ng - 1, j), k);
        }
    } else if (a.kind == 2) {  // inflow left, zero gradient elsewhere
        if (i < ng) {
            for (int k = 0; k < 4; k++) v[k] = a.inflow[k];
        } else {
            int si = i >= nx + ng ? nx + ng - 1 : i;
            int sj = j < ng ? ng : (j >= ny + ng ? ny + ng - 1 : j);
            long long s = L.at(si, sj);
            const double* S = (a.unew_nf && a.type[s] != 0) ? a.unew_nf : U;
            for (int k = 0; k < 4; k++) v[k] = S[s + k * PL];
        }
    } else {
        return;  // the reference policy is written by the step kernel
    }
    long long o = L.at(i, j);
    for (int k = 0; k < 4; k++) U[o + k * PL] = v[k];
}
__global__ void k_edge(EdgeArgs a) { edge_one(a, blockIdx.x * blockDim.x + threadIdx.x); }
// Ghost fill phase 1 and the edge ring in one launch (independent work: the fill
// reads ring donors from the previous state): blocks below nfb fill, the rest
// write the ring.
__global__ __launch_bounds__(FILL_THREADS) void k_fill_edge(FillArgs f, EdgeArgs e, int nfb) {
    if ((int)blockIdx.x < nfb) fill_one(f, blockIdx.x * FILL_THREADS + threadIdx.x);
    else edge_one(e, (blockIdx.x - nfb) * FILL_THREADS + threadIdx.x);
}
// Batched edge ring: block b covers ring cells map[b].y .. +EDGE_THREADS-1 of solve map[b].x.
__global__ void k_edge_batch(const EdgeArgs* __restrict__ slots, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x >= 0) edge_one(slots[m.x], m.y + threadIdx.x);
}

// ------------------------------------------------------ persistent batched step
// Grid barrier among all blocks of a launch (every block resident): arrivals
// counted in bar[0], which only grows during a launch, so barrier k completes
// when it reaches (k + 1) x blocks. Device-scope fences before the arrival
// (release: this block's stores reach the coherence point past its XCD's L2)
// and after the wait (acquire: no stale lines of other XCDs' stores). A wait
// longer than spin_limit polls sets *err and every block leaves the launch.
__device__ __forceinline__ bool grid_sync(const PersistArgs& p, unsigned int& target) {
    target += gridDim.x;
    int ok = 1;
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(p.bar, 1u);
        long long spins = 0;
        while (__hip_atomic_load(p.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > p.spin_limit || __hip_atomic_load(p.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                atomicExch(p.err, 1);
                ok = 0;
                break;
            }
        }
        __threadfence();
    }
    return __shfl(ok, 0, 64) != 0;
}

template <int SCH, int RS, int LIM, typename T>
__global__ __launch_bounds__(64) void k_persist(PersistArgs p) {
    const int G = gridDim.x;
    unsigned int target = 0;
    const int nstrip = p.counts[0], nghost = p.counts[1], nring = p.counts[2];
    for (int s = 0; s < p.nsteps; s++) {
        for (int stage = p.rk2 ? 1 : 0;;) {
            // strips of every resident solve (TimeAdvance; reference ring fused)
            for (int w = blockIdx.x; w < nstrip; w += G) {
                const int2 m = p.smap[w];
                const StepArgs& a = p.step[m.x];
                step_strip<SCH, RS, LIM, T, false>(a, m.y % a.nbx, m.y / a.nbx, stage);
            }
            if (!grid_sync(p, target)) return;
            // ComputeDt / mass / clock per solve (RK2 stage 1: make the stage buffer current)
            for (int q = blockIdx.x; q < p.nslots; q += G)
                if (p.fin[q].ctl) finish_body<64>(p.fin[q], stage == 1 ? 2 : 1);
            if (!grid_sync(p, target)) return;
            // Boundary phase 1 and Global_Boundary
            if (p.has_fill)
                for (int w = blockIdx.x; w < nghost; w += G) {
                    const int2 m = p.gmap[w];
                    for (int k = threadIdx.x; k < FILL_THREADS; k += 64) fill_one(p.fill[m.x], m.y + k);
                }
            for (int w = blockIdx.x; w < nring; w += G) {
                const int2 m = p.rmap[w];
                for (int k = threadIdx.x; k < EDGE_THREADS; k += 64) edge_one(p.edge[m.x], m.y + k);
            }
            if (!grid_sync(p, target)) return;
            // Boundary phase 2
            if (p.has_fill) {
                for (int w = blockIdx.x; w < nghost; w += G) {
                    const int2 m = p.gmap[w];
                    for (int k = threadIdx.x; k < FILL_THREADS; k += 64) scatter_one(p.fill[m.x], m.y + k);
                }
                if (!grid_sync(p, target)) return;
            }
            if (stage == 1) { stage = 2; continue; }
            break;
        }
    }
}

__global__ void k_zero_u32(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Reference Global_Boundary as a separate pass (used with Lax, whose step
// kernel does not write the ring): reads U_NEW = the new state (ring of U_NEW
// taken from the previous state for the corners), writes the ring.
__global__ void k_edge_ref(EdgeArgs a, const double* corner, int corner_from_input) {
    if (a.require_ran && !a.ctl->ran) return;
    const Layout L = a.L;
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int nring = L.M * L.N - nx * ny;
    if (q >= nring) return;
    int i, j;
    int nb = L.M * ng;
    if (q < nb) { j = q / L.M; i = q % L.M; }
    else if (q < 2 * nb) { int r = q - nb; j = ny + ng + r / L.M; i = r % L.M; }
    else { int r = q - 2 * nb; j = ng + r / (2 * ng); int s = r % (2 * ng); i = s < ng ? s : nx + s; }
    const int cur = a.ctl->cur;
    double* U = a.buf[cur];
    const double* Uprev = a.buf[cur ^ 1];
    const long long PL = L.plane;
    const long long o = L.at(i, j);
    if (i >= ng && i < nx + ng) {
        long long s = (j < ng) ? L.at(i, 2 * ng - 1) : L.at(i, ny + ng - 1);
        long long s2 = (j < ng) ? L.at(i, 2 * ng - j - 1) : L.at(i, 2 * (ny + ng) - 1 - j);
        U[o] = U[s]; U[o + PL] = U[s + PL]; U[o + 3 * PL] = U[s + 3 * PL];
        U[o + 2 * PL] = -U[s2 + 2 * PL];
    } else {
        int scol = i < ng ? ng : nx + ng - 1;
        bool ringrow = j < ng || j >= ny + ng;
        if (!ringrow) {
            long long s = L.at(scol, j);
            for (int k = 0; k < 4; k++) U[o + k * PL] = U[s + k * PL];
        } else {
            int side = i < ng ? 0 : 1, ri = j < ng ? j : j - ny;
            for (int k = 0; k < 4; k++)
                U[o + k * PL] = corner_from_input ? Uprev[L.at(scol, j) + k * PL] : corner[(side * 2 * ng + ri) * 4 + k];
        }
    }
}

// ------------------------------------------------------------------ launchers
// A batched launch (k_step_batch) instead of a single-solve one when non-null.
struct BatchLaunch {
    const StepArgs* slots;
    const int2* map;
    int nblocks;
    int stage;
};
template <int SCH, int RS, int LIM>
static hipError_t go(int prec, const StepArgs* a, const BatchLaunch* b, hipStream_t s) {
    if (b) {
        if (prec == 1)
            hipLaunchKernelGGL((k_step_batch<SCH, RS, LIM, float>), dim3(b->nblocks), dim3(64), 0, s, b->slots, b->map,
                               b->stage);
        else
            hipLaunchKernelGGL((k_step_batch<SCH, RS, LIM, double>), dim3(b->nblocks), dim3(64), 0, s, b->slots,
                               b->map, b->stage);
    } else if (a->cut) {  // cut-cell variant (its extra registers stay out of the plain kernel)
        if (prec == 1) hipLaunchKernelGGL((k_step<SCH, RS, LIM, float, true>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
        else hipLaunchKernelGGL((k_step<SCH, RS, LIM, double, true>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
    } else if (a->edge_fused == 2) {  // periodic ring written by the step kernel
        if (prec == 1) hipLaunchKernelGGL((k_step_periodic<SCH, RS, LIM, float>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
        else hipLaunchKernelGGL((k_step_periodic<SCH, RS, LIM, double>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
    } else {
        if (prec == 1) hipLaunchKernelGGL((k_step<SCH, RS, LIM, float, false>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
        else hipLaunchKernelGGL((k_step<SCH, RS, LIM, double, false>), dim3(a->nbx, a->nby), dim3(64), 0, s, *a);
    }
    return hipGetLastError();
}
template <int SCH, int RS>
static hipError_t go_l(int lim, int prec, const StepArgs* a, const BatchLaunch* b, hipStream_t s) {
    if (lim == LIM_MINMOD) return go<SCH, RS, LIM_MINMOD>(prec, a, b, s);
    if (lim == LIM_VANLEER) return go<SCH, RS, LIM_VANLEER>(prec, a, b, s);
    return go<SCH, RS, LIM_VANALBADA>(prec, a, b, s);
}
template <int SCH>
static hipError_t go_r(int rs, int lim, int prec, const StepArgs* a, const BatchLaunch* b, hipStream_t s) {
    if (rs == RS_SW) return go_l<SCH, RS_SW>(lim, prec, a, b, s);
    if (rs == RS_PVRS) return go_l<SCH, RS_PVRS>(lim, prec, a, b, s);
    return go_l<SCH, RS_HLL>(lim, prec, a, b, s);
}
static hipError_t go_s(int scheme, int rs, int lim, int prec, const StepArgs* a, const BatchLaunch* b,
                       hipStream_t s) {
    if (scheme == SCH_MHM) return go_r<SCH_MHM>(rs, lim, prec, a, b, s);
    if (scheme == SCH_GODUNOV) return go_r<SCH_GODUNOV>(rs, LIM_MINMOD, prec, a, b, s);
    return go_r<SCH_MUSCL>(rs, lim, prec, a, b, s);
}
hipError_t launch_step(int scheme, int rs, int lim, int prec, const StepArgs& a, hipStream_t s) {
    if (scheme == SCH_LAX) return launch_step_lax(a, s);
    if (a.rowpar) return launch_step_rowpar(scheme, rs, lim, prec, a, s);
    return go_s(scheme, rs, lim, prec, &a, nullptr, s);
}
hipError_t launch_step_batch(int scheme, int rs, int lim, int prec, const StepArgs* slots, const int2* map,
                             int nblocks, int stage, hipStream_t s) {
    if (scheme == SCH_LAX) return hipErrorInvalidValue;
    BatchLaunch b{slots, map, nblocks, stage};
    return go_s(scheme, rs, lim, prec, nullptr, &b, s);
}
// persistent batched step: the kernel for a slot selection, and its resident grid
template <int SCH, int RS, int LIM>
static const void* persist_fn(int prec) {
    if (prec == IBTFO_MIXED) return (const void*)k_persist<SCH, RS, LIM, float>;
    return (const void*)k_persist<SCH, RS, LIM, double>;
}
static const void* persist_kernel(int scheme, int rs, int lim, int prec) {
    auto by_lim = [&](auto sch, auto r) -> const void* {
        constexpr int S = decltype(sch)::value, R = decltype(r)::value;
        if (lim == LIM_MINMOD) return persist_fn<S, R, LIM_MINMOD>(prec);
        if (lim == LIM_VANLEER) return persist_fn<S, R, LIM_VANLEER>(prec);
        return persist_fn<S, R, LIM_VANALBADA>(prec);
    };
    auto by_rs = [&](auto sch) -> const void* {
        if (rs == RS_SW) return by_lim(sch, std::integral_constant<int, RS_SW>());
        if (rs == RS_PVRS) return by_lim(sch, std::integral_constant<int, RS_PVRS>());
        return by_lim(sch, std::integral_constant<int, RS_HLL>());
    };
    if (scheme == SCH_MHM) return by_rs(std::integral_constant<int, SCH_MHM>());
    if (scheme == SCH_GODUNOV) return by_rs(std::integral_constant<int, SCH_GODUNOV>());
    return by_rs(std::integral_constant<int, SCH_MUSCL>());
}
int persist_blocks(int scheme, int rs, int lim, int prec) {
    int per_cu = 0, dev = 0, cus = 0;
    if (scheme == SCH_LAX) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel(scheme, rs, lim, prec), 64, 0) !=
            hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}
hipError_t launch_persist(int scheme, int rs, int lim, int prec, const PersistArgs& p, int blocks, hipStream_t s) {
    if (scheme == SCH_LAX || blocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, p.bar, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    void* args[] = {(void*)&p};
    return hipLaunchKernel(persist_kernel(scheme, rs, lim, prec), dim3(blocks), dim3(64), args, 0, s);
}
hipError_t launch_finish_batch(const FinishArgs* f, int nslots, int advance, hipStream_t s) {
    hipLaunchKernelGGL(k_finish_batch, dim3(nslots), dim3(FINISH_THREADS), 0, s, f, nslots, advance);
    return hipGetLastError();
}
hipError_t launch_fill_batch(const FillArgs* f, const int2* map, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_batch, dim3(nblocks), dim3(FILL_THREADS), 0, s, f, map);
    return hipGetLastError();
}
hipError_t launch_scatter_batch(const FillArgs* f, const int2* map, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_batch, dim3(nblocks), dim3(FILL_THREADS), 0, s, f, map);
    return hipGetLastError();
}
hipError_t launch_edge_batch(const EdgeArgs* e, const int2* map, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_edge_batch, dim3(nblocks), dim3(EDGE_THREADS), 0, s, e, map);
    return hipGetLastError();
}
// Batched status read-out: one record per slot (idle slots: ctl null -> zeros).
__global__ void k_gather(const FinishArgs* __restrict__ f, int nslots, SlotStatus* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    SlotStatus st{};
    const Ctl* c = f[i].ctl;
    if (c) {
        st.iter = c->iter; st.done = c->done; st.status = c->status;
        st.now_t = c->now_t; st.dt = c->dt; st.mass = c->mass_now;
    }
    out[i] = st;
}
hipError_t launch_gather(const FinishArgs* f, int nslots, SlotStatus* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3((nslots + 63) / 64), dim3(64), 0, s, f, nslots, out);
    return hipGetLastError();
}
hipError_t launch_step_lax(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_step_lax, dim3(a.nbx, a.nby), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_fill(const FillArgs& a, hipStream_t s) {
    if (a.nghost == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((a.nghost + FILL_THREADS - 1) / FILL_THREADS), dim3(FILL_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_scatter(const FillArgs& a, hipStream_t s) {
    if (a.nghost == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((a.nghost + FILL_THREADS - 1) / FILL_THREADS), dim3(FILL_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_fill_edge(const FillArgs& f, const EdgeArgs& e, hipStream_t s) {
    if (f.nghost == 0) return launch_edge(e, s);
    const int nfb = (f.nghost + FILL_THREADS - 1) / FILL_THREADS;
    const int nring = e.L.M * e.L.N - e.L.nx * e.L.ny;
    const int neb = (nring + FILL_THREADS - 1) / FILL_THREADS;
    hipLaunchKernelGGL(k_fill_edge, dim3(nfb + neb), dim3(FILL_THREADS), 0, s, f, e, nfb);
    return hipGetLastError();
}
hipError_t launch_edge(const EdgeArgs& a, hipStream_t s) {
    int nring = a.L.M * a.L.N - a.L.nx * a.L.ny;
    hipLaunchKernelGGL(k_edge, dim3((nring + EDGE_THREADS - 1) / EDGE_THREADS), dim3(EDGE_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_edge_ref(const EdgeArgs& a, const double* corner, int corner_from_input, hipStream_t s) {
    int nring = a.L.M * a.L.N - a.L.nx * a.L.ny;
    hipLaunchKernelGGL(k_edge_ref, dim3((nring + 255) / 256), dim3(256), 0, s, a, corner, corner_from_input);
    return hipGetLastError();
}
// A uniform initial state written on the device into up to three state
// buffers (every element of each of the 4 planes); block 0 also makes buffer 0
// current and clears the step flag, as an upload of the state does.
__global__ void k_uniform_state(double* s0, double* s1, double* s2, long long plane, double u0, double u1,
                                double u2, double u3, Ctl* ctl) {
    const long long n = 4 * plane;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane);
        const double v = k == 0 ? u0 : (k == 1 ? u1 : (k == 2 ? u2 : u3));
        s0[q] = v;
        s1[q] = v;
        if (s2) s2[q] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->cur = 0;
        ctl->ran = 0;
    }
}
hipError_t launch_uniform_state(double* s0, double* s1, double* s2, long long plane, const double* U, Ctl* ctl,
                                hipStream_t s) {
    const long long n = 4 * plane;
    const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_uniform_state, dim3(blocks), dim3(256), 0, s, s0, s1, s2, plane, U[0], U[1], U[2], U[3], ctl);
    return hipGetLastError();
}

hipError_t launch_merge(const MergeArgs& a, hipStream_t s) {
    if (a.ngroups == 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge, dim3(merge_blocks(a.ngroups)), dim3(MERGE_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(a.nblocks), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FinishArgs f;
    f.ctl = a.ctl; f.part = a.mass_part; f.nblocks = a.nblocks;
    f.cfl = a.cfl; f.dx = a.dx; f.t_end = a.t_end; f.max_iter = a.max_iter;
    f.advance = 0; f.reset = a.reset;
    return launch_finish(f, s);
}
hipError_t launch_finish(const FinishArgs& f, hipStream_t s) {
    const int nrec = f.nblocks + (f.part2 ? f.nblocks2 : 0);
    if (f.advance == 1 && f.fold && nrec >= FOLD_MIN_PARTS) {
        FinishArgs g = f;
        g.nfold = std::min(FOLD_MAX, (nrec + FOLD_THREADS - 1) / FOLD_THREADS);
        g.fold_per = (nrec + g.nfold - 1) / g.nfold;
        hipLaunchKernelGGL(k_fold, dim3(g.nfold), dim3(FOLD_THREADS), 0, s, g);
        return hipGetLastError();
    }
    if (nrec <= FINISH_SMALL_MAX) hipLaunchKernelGGL(k_finish_small, dim3(1), dim3(64), 0, s, f);
    else hipLaunchKernelGGL(k_finish, dim3(1), dim3(FINISH_THREADS), 0, s, f);
    return hipGetLastError();
}

}  // namespace ibtfo
