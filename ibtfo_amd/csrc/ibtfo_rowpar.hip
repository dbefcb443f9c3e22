// Row-parallel step kernel (k_step_rowpar): the small-grid form of the step
// kernel, each row of a strip on its own wave. It lives in a translation unit
// of its own: compiled beside the marching strip kernels, its instantiations
// changed their register allocation (mixed-precision variants from 124 to 131
// VGPRs, four waves per SIMD to three).
#include "ibtfo_internal.h"
#include "ibtfo_face.h"

namespace ibtfo {

// The update of cell (c, j) from its face fluxes (F, G on its upper x / y faces;
// fm_lds[k * 64], gm_lds[k * 64] on its lower faces, in the LDS slots of the
// lanes that computed them): the conservative or cut-cell update, floors,
// dt/mass partial terms, the RK2 stage average, the store and the fused
// Global_Boundary ring writes. The same statements as the update inside
// step_strip, which keeps its own inline copy: factoring it out of the strip
// kernel changed that kernel's register allocation (the mixed-precision
// cut-cell variants went from 128 to 165 VGPRs, 4 to 3 waves per SIMD). The
// two copies are held equal by the row-parallel tests, which compare the two
// kernels' states bit for bit over every scheme, flux, boundary and stage.
template <int SCH, int RS, int LIM, typename T, bool CUT>
__device__ __forceinline__ void update_cell(const StepArgs& a, const Layout& L, const int c, const int j,
                                            const double (&w0)[4], const T (&F)[4], const T (&G)[4],
                                            const T* fm_lds, const T* gm_lds, const double dt,
                                            const Rcp<double>& r_dx, const Rcp<double>& r_dy,
                                            const Rcp<double>& r_gm1, const double dtdx, const double dtdy,
                                            double* __restrict__ Uo, const int stage, double& amax, double& umax,
                                            double& vmax, double& msum, int& bad) {
    constexpr bool FP64 = sizeof(T) == sizeof(double);
    const long long PL = L.plane;
    const double dx = a.dx, dy = a.dy;
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    double un[4];
    const long long o = L.at(c, j);
    const uint8_t tc = a.type[o];
    const bool fluid = CUT ? (tc & 3) == 0 : tc == 0;
    const bool merged = CUT && (tc & MERGED_BIT);
    if (fluid) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (FP64) {
                double Fm = fm_lds[k * 64];
                double v = w0[k] - dq(dt * ((double)F[k] - Fm), r_dx);
                un[k] = v - dq(dt * ((double)G[k] - (double)gm_lds[k * 64]), r_dy);
            } else {  // FP32 fluxes, FP64 accumulation
                double Fm = (double)fm_lds[k * 64];
                un[k] = w0[k] - dtdx * ((double)F[k] - Fm) - dtdy * ((double)G[k] - (double)gm_lds[k * 64]);
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
        const double rho = w0[0], u0 = w0[1] / rho, v0 = w0[2] / rho;
        const double pw = a.gm1 * (w0[3] - 0.5 * rho * (u0 * u0 + v0 * v0));
        const double wall[4] = {0.0, pw * cc.sbx, pw * cc.sby, 0.0};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double Fk = (double)F[k], Fm = (double)fm_lds[k * 64];
            const double Gk = (double)G[k], Gm = (double)gm_lds[k * 64];
            double v = w0[k] - dt * ((cc.ae * Fk - cc.aw * Fm) * dy) / (cc.alpha * dx * dy);
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
        for (int k = 0; k < 4; k++) un[k] = w0[k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) Uo[o + k * PL] = un[k];
    if (a.edge_fused == 2) periodic_images(Uo, L, c, j, un);
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

// Deferred finish: the dt / mass reduction of a step's nrec <= DF_MAX_REC
// partial records by one wave, in the order of the one-wave finish kernel
// (k_finish_small: lane t takes records t, t + 64, ... in turn, then the
// cross-lane max / sum), so dt, mass and the clock come out bit for bit as
// they do with a finish kernel between the steps.
constexpr int DF_PER_LANE = DF_MAX_REC / 64;
__device__ __forceinline__ void df_reduce(const double* part, const int nrec, double& A, double& Um, double& Vm,
                                          double& m, int& bad) {
    const int t = threadIdx.x & 63;
    double ra[DF_PER_LANE], ru[DF_PER_LANE], rv[DF_PER_LANE], rm[DF_PER_LANE], rb[DF_PER_LANE];
#pragma unroll
    for (int b = 0; b < DF_PER_LANE; b++) {
        const int q = t + 64 * b;
        const double* r = part + PART_STRIDE * (long long)(q < nrec ? q : 0);
        ra[b] = r[0]; ru[b] = r[1]; rv[b] = r[2]; rm[b] = r[3]; rb[b] = r[4];
    }
    A = 0.0; Um = 0.0; Vm = 0.0; m = 0.0;
    bad = 0;
#pragma unroll
    for (int b = 0; b < DF_PER_LANE; b++) {
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
}

// The finish of the step that ran with head `h` (the finish kernel's clock
// bookkeeping): dt from the reduced partials, the reference's clock order
// (now_t advances by the new dt), the log entry of that step by `log_lane`.
__device__ __forceinline__ DfHead df_advance(const StepArgs& a, const DfHead& h, double A2, double Um, double Vm,
                                             double m, int bad, bool log_lane) {
    const double A = ib_sqrt(A2);
    const double dt = a.cfl * a.dx / (A + ((Um < Vm) ? Vm : Um));
    DfHead n;
    n.dt = dt;
    n.now_t = h.now_t + dt;
    n.iter = h.iter + 1;
    n.mass = m;
    n.done = !(n.now_t < a.t_end && n.iter < a.max_iter);
    n.cur = h.cur ^ 1;
    n.pending = 0;
    if (log_lane) {
        Ctl* ctl = a.ctl;
        const int slot = h.iter % LOGCAP;
        ctl->log_dt[slot] = dt;
        ctl->log_t[slot] = n.now_t;
        ctl->log_mass[slot] = m;
        if (bad) ctl->status |= 1;
    }
    return n;
}

// The head of this launch: from the control block for the first kernel of a
// batch, else the previous kernel's head advanced over its partial records.
// Every wave computes it alike; block (0,0) stores it for the next kernel.
__device__ __forceinline__ DfHead df_head(const StepArgs& a, const bool first_block) {
    Ctl* ctl = a.ctl;
    DfHead h;
    if (a.df == 1) {
        h.dt = ctl->dt; h.now_t = ctl->now_t; h.mass = ctl->mass_now;
        h.iter = ctl->iter; h.done = ctl->done; h.cur = ctl->cur; h.pending = 0;
    } else {
        const int p = a.df_k ^ 1;
        const DfHead prev = ctl->df[p];
        double A, Um, Vm, m;
        int bad;
        df_reduce(a.df_part + (long long)p * DF_MAX_REC * PART_STRIDE, a.nbx * a.nby, A, Um, Vm, m, bad);
        h = prev.pending ? df_advance(a, prev, A, Um, Vm, m, bad, first_block && threadIdx.x == 0) : prev;
    }
    h.pending = !h.done;
    if (first_block && threadIdx.x == 0) ctl->df[a.df_k] = h;
    return h;
}

// Row-parallel strip (small grids, where a strip's rows marching one after the
// other set the step's latency): the block has rows + 1 waves over the same
// 60-column x `rows` strip as step_strip; wave w takes row j = j0 - 1 + w at
// once. It loads rows j-1 .. j+2 of its column, reconstructs them, computes
// the x faces of row j (for j >= j0, or in quirks mode where the y face may
// take the aliased x flux) and the y face j+1/2, and after a block barrier the
// waves of rows j0 .. j1-1 update their cells with the y flux of face j-1/2
// from the wave below. Every face is computed by the same statements as in
// step_strip, so the state is bit for bit the same; only the grouping of the
// mass partial sum differs. Block barriers are reached by every wave.
template <int SCH, int RS, int LIM, typename T, bool CUT>
__device__ __forceinline__ void step_rowpar(const StepArgs& a, const int bx, const int by, const int stage) {
    Ctl* ctl = a.ctl;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const Layout L = a.L;
    const long long PL = L.plane;
    const int ng = L.ng, nx = L.nx, ny = L.ny;
    const int cb = ng + STRIP_W * bx;
    const int c = cb - 2 + lane;
    const int cl = min(max(c, 0), L.M - 1);
    const int j0 = ng + a.rows * by;
    const int j1 = min(j0 + a.rows, ny + ng);
    const int j = j0 - 1 + wv;
    const bool live = j < j1;  // the waves past a short last strip only reach the barriers

    // rows j-1 .. j+2 of column c (in range for every live wave: ng >= 2) of both
    // state buffers, read together with the control block: the rows' addresses
    // do not depend on which buffer is current, so no load waits for another
    // (on a latency-bound small grid that is one memory round trip less)
    double q0[4][4], q1[4][4];
    if (live) {
        const long long o = L.at(cl, j - 1);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                q0[r][k] = a.buf[0][o + r * L.pitch + k * PL];
                q1[r][k] = a.buf[1][o + r * L.pitch + k * PL];
            }
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int k = 0; k < 4; k++) q0[r][k] = q1[r][k] = 0.0;
    }
    int cur;
    double dt;
    if (a.df) {
        const DfHead h = df_head(a, bx == 0 && by == 0);
        if (h.done) return;
        cur = h.cur;
        dt = h.dt;
    } else {
        if (ctl->done) {
            if (bx == 0 && by == 0 && threadIdx.x == 0) ctl->ran = 0;
            return;
        }
        cur = ctl->cur ^ (stage == 3 ? 1 : 0);  // the input buffer
        dt = ctl->dt;
    }
    const double* __restrict__ Ui = a.buf[cur];
    double* __restrict__ Uo = a.buf[cur ^ 1];
    GasT<T> g;
    g.gam = (T)a.gam; g.gm1 = (T)a.gm1; g.gp1 = (T)a.gp1; g.c1 = (T)a.c1;
    g.r_gam = Rcp<T>{(T)a.gam, (T)a.r_gam};
    g.r_gm1 = Rcp<T>{(T)a.gm1, (T)a.r_gm1};
    g.r_dx = Rcp<T>{(T)a.dx, (T)a.r_dx};
    g.r_dy = Rcp<T>{(T)a.dy, (T)a.r_dy};
    const T dtc = (T)dt;
    const Rcp<double> r_dx{a.dx, a.r_dx}, r_dy{a.dy, a.r_dy}, r_gm1{a.gm1, a.r_gm1};
    const double dtdx = dt * a.r_dx, dtdy = dt * a.r_dy;  // mixed form only

    __shared__ double rb[ROWPAR_MAX_ROWS + 1][4][64];
    __shared__ double hb[ROWPAR_MAX_ROWS + 1][4][64];
    __shared__ T fb[ROWPAR_MAX_ROWS + 1][4][64];
    __shared__ T gb[ROWPAR_MAX_ROWS + 1][4][64];
    __shared__ double sp[ROWPAR_MAX_ROWS + 1][5];

    // reference Global_Boundary corners (Test/Sod/Global_Boudary.cpp:62-75), as in step_strip
    if (a.edge_fused == 1 && bx == 0 && by == 0 && wv == 0 && lane < 4 * ng * ng) {
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

    double wm[4], w0[4], w1[4], w2[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        wm[k] = cur ? q1[0][k] : q0[0][k];
        w0[k] = cur ? q1[1][k] : q0[1][k];
        w1[k] = cur ? q1[2][k] : q0[2][k];
        w2[k] = cur ? q1[3][k] : q0[3][k];
    }
    const bool xrow = live && (j >= j0 || (a.quirks && RS != RS_SW));
    const int ln = min(lane + 1, 63), lp = max(lane - 1, 0);
    T F[4] = {0, 0, 0, 0};
    double xp[4], xlo[4];
#pragma unroll
    for (int k = 0; k < 4; k++) rb[wv][k][lane] = w0[k];
    __syncthreads();
    if (xrow) {
        double xm[4], xhi[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xm[k] = rb[wv][k][lp];
            xp[k] = rb[wv][k][ln];
        }
        cell_states<SCH, LIM, T>(xm, w0, xp, xlo, xhi, a.quirks);
#pragma unroll
        for (int k = 0; k < 4; k++) hb[wv][k][lane] = xhi[k];
    }
    __syncthreads();
    if (xrow) {
        double xr[4];
#pragma unroll
        for (int k = 0; k < 4; k++) xr[k] = hb[wv][k][ln];
        xface<SCH, RS>(g, a.gm1, dtc, a.quirks, w0, xp, xlo, xr, F);
    }
    T G[4] = {0, 0, 0, 0};
    if (live) {
        double ylo[4], hi_unused[4], ynlo[4], ynhi[4];
        cell_states<SCH, LIM, T>(wm, w0, w1, ylo, hi_unused, a.quirks);
        cell_states<SCH, LIM, T>(w0, w1, w2, ynlo, ynhi, a.quirks);
        yface<SCH, RS>(g, a.gm1, dtc, a.quirks, w0, w1, ylo, ynhi, F, G);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fb[wv][k][lane] = F[k];
        gb[wv][k][lane] = G[k];
    }
    __syncthreads();
    double amax = 0.0, umax = 0.0, vmax = 0.0, msum = 0.0;
    int bad = 0;
    if (live && j >= j0 && lane >= 2 && lane < STRIP_W + 2 && c < nx + ng)
        update_cell<SCH, RS, LIM, T, CUT>(a, L, c, j, w0, F, G, &fb[wv][0][lane - 1], &gb[wv - 1][0][lane], dt,
                                          r_dx, r_dy, r_gm1, dtdx, dtdy, Uo, stage, amax, umax, vmax, msum, bad);
    // one partial record per block: each wave's, then the waves in order
    amax = wave_max(amax);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    msum = wave_sum(msum);
    bad = __any(bad);
    if (lane == 0) {
        sp[wv][0] = amax; sp[wv][1] = umax; sp[wv][2] = vmax; sp[wv][3] = msum; sp[wv][4] = bad ? 1.0 : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        double A = sp[0][0], U = sp[0][1], V = sp[0][2], m = sp[0][3], b = sp[0][4];
        for (int q = 1; q < nw; q++) {
            A = (A < sp[q][0]) ? sp[q][0] : A;
            U = (U < sp[q][1]) ? sp[q][1] : U;
            V = (V < sp[q][2]) ? sp[q][2] : V;
            m = m + sp[q][3];
            b = (b < sp[q][4]) ? sp[q][4] : b;
        }
        double* part = a.df ? a.df_part + (long long)a.df_k * DF_MAX_REC * PART_STRIDE : a.mass_part;
        double* r = part + PART_STRIDE * (long long)(by * a.nbx + bx);
        r[0] = A; r[1] = U; r[2] = V; r[3] = m; r[4] = b;
    }
}

// The end of a deferred-finish batch: the finish of its last kernel (parity
// k), if that kernel ran its step, published to the control block as the
// finish kernel leaves it (ran = whether that step ran).
__global__ __launch_bounds__(64) void k_df_flush(StepArgs a, int k) {
    Ctl* ctl = a.ctl;
    const DfHead h = ctl->df[k];
    double A, Um, Vm, m;
    int bad;
    df_reduce(a.df_part + (long long)k * DF_MAX_REC * PART_STRIDE, a.nbx * a.nby, A, Um, Vm, m, bad);
    if (threadIdx.x != 0) return;
    const DfHead n = h.pending ? df_advance(a, h, A, Um, Vm, m, bad, true) : h;
    ctl->dt = n.dt;
    ctl->now_t = n.now_t;
    ctl->iter = n.iter;
    ctl->done = n.done;
    ctl->cur = n.cur;
    ctl->mass_now = n.mass;
    ctl->ran = h.pending;
}

template <int SCH, int RS, int LIM, typename T, bool CUT>
__global__ __launch_bounds__(64 * (ROWPAR_MAX_ROWS + 1)) void k_step_rowpar(StepArgs a) {
    step_rowpar<SCH, RS, LIM, T, CUT>(a, blockIdx.x, blockIdx.y, a.stage);
}

template <int SCH, int RS, int LIM>
static hipError_t go_rowpar(int prec, const StepArgs& a, hipStream_t s) {
    const dim3 grid(a.nbx, a.nby), block(64 * (a.rows + 1));
    if (a.cut) {
        if (prec == 1) hipLaunchKernelGGL((k_step_rowpar<SCH, RS, LIM, float, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step_rowpar<SCH, RS, LIM, double, true>), grid, block, 0, s, a);
    } else {
        if (prec == 1) hipLaunchKernelGGL((k_step_rowpar<SCH, RS, LIM, float, false>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step_rowpar<SCH, RS, LIM, double, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}
template <int SCH, int RS>
static hipError_t go_rowpar_l(int lim, int prec, const StepArgs& a, hipStream_t s) {
    if (lim == LIM_MINMOD) return go_rowpar<SCH, RS, LIM_MINMOD>(prec, a, s);
    if (lim == LIM_VANLEER) return go_rowpar<SCH, RS, LIM_VANLEER>(prec, a, s);
    return go_rowpar<SCH, RS, LIM_VANALBADA>(prec, a, s);
}
template <int SCH>
static hipError_t go_rowpar_r(int rs, int lim, int prec, const StepArgs& a, hipStream_t s) {
    if (rs == RS_SW) return go_rowpar_l<SCH, RS_SW>(lim, prec, a, s);
    if (rs == RS_PVRS) return go_rowpar_l<SCH, RS_PVRS>(lim, prec, a, s);
    return go_rowpar_l<SCH, RS_HLL>(lim, prec, a, s);
}

hipError_t launch_df_flush(const StepArgs& a, int k, hipStream_t s) {
    if (!a.df_ok || a.nbx * a.nby > DF_MAX_REC || (k != 0 && k != 1)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_df_flush, dim3(1), dim3(64), 0, s, a, k);
    return hipGetLastError();
}

hipError_t launch_step_rowpar(int scheme, int rs, int lim, int prec, const StepArgs& a, hipStream_t s) {
    if (scheme == SCH_LAX || a.rows < 1 || a.rows > ROWPAR_MAX_ROWS) return hipErrorInvalidValue;
    if (a.df && (!a.df_ok || a.nbx * a.nby > DF_MAX_REC || a.stage != 0 || !a.df_part)) return hipErrorInvalidValue;
    if (scheme == SCH_MHM) return go_rowpar_r<SCH_MHM>(rs, lim, prec, a, s);
    if (scheme == SCH_GODUNOV) return go_rowpar_r<SCH_GODUNOV>(rs, lim, prec, a, s);
    return go_rowpar_r<SCH_MUSCL>(rs, lim, prec, a, s);
}

}  // namespace ibtfo
