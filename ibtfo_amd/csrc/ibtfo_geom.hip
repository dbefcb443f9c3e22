// Geometry kernels: cell centres, level-set hooks, normals, fluid/solid/ghost
// classification and the ghost-fluid donor records, run once per solve.
//
// All of it is integer logic and IEEE +,-,*,/,sqrt in the reference's order
// (no contraction), so flags, mirror points, donor indices and weights are bit
// identical to the reference's Src/Geometry and Scheme_Index results.
#include "ibtfo_geom.h"

namespace ibtfo {

// Psy_coord.cpp:25-35: x_i = Li + i*dx, y_j = Hi + j*dx (the reference spaces y
// with dx as well).
__device__ __forceinline__ double xc(const GeomArgs& g, int i) { return g.Li + i * g.dx; }
__device__ __forceinline__ double yc(const GeomArgs& g, int j) { return g.Hi + j * g.dx; }

__device__ double hook_phi(const GeomArgs& g, double x, double y, long long q) {
    switch (g.kind) {
        case 1: {  // Test/InclineSod/Level_Set_function.cpp:8-20, :48-51
            double k = g.p[0], nrm = g.p[1], b1 = g.p[2], b2 = g.p[3];
            double p1 = (y > k * x + b1) ? 1 * ib_fabs(k * x - y + b1) / nrm : -1 * ib_fabs(k * x - y + b1) / nrm;
            double p2 = (y > k * x + b2) ? -1 * ib_fabs(k * x - y + b2) / nrm : 1 * ib_fabs(k * x - y + b2) / nrm;
            return (p1 < p2) ? p1 : p2;
        }
        case 2: {  // Test/Vortex/Level_Set_function.cpp:31-38
            double r = ib_sqrt((x - 0.50) * (x - 0.50) + (y - 0.50) * (y - 0.50));
            double p1 = g.p[1] - r, p2 = r - g.p[0];
            return (p1 < p2) ? p1 : p2;
        }
        case 3:
            return g.phi_in[q];
        case 4: {  // ellipse body: phi > 0 outside; |phi| approximates the distance
            double cx = g.p[0], cy = g.p[1], a = g.p[2], b = g.p[3], ca = g.p[4], sa = g.p[5];
            double dxp = x - cx, dyp = y - cy;
            double xr = dxp * ca + dyp * sa, yr = -dxp * sa + dyp * ca;
            double s = ib_sqrt((xr / a) * (xr / a) + (yr / b) * (yr / b));
            // first-order distance: (s - 1) / |grad s|
            double gx = xr / (a * a), gy = yr / (b * b);
            double gn = ib_sqrt(gx * gx + gy * gy);
            return (s > 0.0 && gn > 0.0) ? (s * s - 1.0) / (2.0 * gn) : -((a < b) ? a : b);
        }
        case 5: {  // union of circles and convex polygons; exact signed distance (< 0 inside a body)
            double best = 1e300;
            for (int k = 0; k < g.nshapes; k++) {
                const double* sh = g.shapes + k * SHAPE_STRIDE;
                double d;
                if (sh[0] == 0.0) {
                    d = ib_sqrt((x - sh[1]) * (x - sh[1]) + (y - sh[2]) * (y - sh[2])) - sh[3];
                } else {
                    int nv = (int)sh[4];
                    double dmin = 1e300;
                    bool inside = true;
                    for (int e = 0; e < nv; e++) {
                        double ax = sh[5 + 2 * e], ay = sh[6 + 2 * e];
                        double bx = sh[5 + 2 * ((e + 1) % nv)], by = sh[6 + 2 * ((e + 1) % nv)];
                        double ex = bx - ax, ey = by - ay, px = x - ax, py = y - ay;
                        double t = (px * ex + py * ey) / (ex * ex + ey * ey);
                        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
                        double qx = px - t * ex, qy = py - t * ey;
                        double dd = qx * qx + qy * qy;
                        dmin = dd < dmin ? dd : dmin;
                        if (ex * py - ey * px < 0.0) inside = false;  // vertices counter-clockwise
                    }
                    d = inside ? -ib_sqrt(dmin) : ib_sqrt(dmin);
                }
                best = d < best ? d : best;
            }
            return best;
        }
        default:
            return 1.0;
    }
}

__global__ void k_geom_phi(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.L.M) return;
    long long o = g.L.at(i, j);
    double x = xc(g, i), y = yc(g, j);
    g.phi[o] = hook_phi(g, x, y, (long long)j * g.L.M + i);
}

// Level_Set.cpp:71-86 (normals, ring left at 0) and :90-158 (classification in
// gather form: a solid cell is a ghost when a fluid cell lies within ng cells
// along x or y; the domain ring is ghost).
__global__ void k_geom_classify(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    if (i >= L.M) return;
    const long long o = L.at(i, j);
    double nxv = 0.0, nyv = 0.0;
    if (i >= 1 && i < L.M - 1 && j >= 1 && j < L.N - 1) {
        nxv = (g.phi[L.at(i + 1, j)] - g.phi[L.at(i - 1, j)]) / ((xc(g, i + 1) - xc(g, i - 1)));
        nyv = (g.phi[L.at(i, j + 1)] - g.phi[L.at(i, j - 1)]) / ((yc(g, j + 1) - yc(g, j - 1)));
    }
    if (g.unit_normals) {  // corrected mode: unit normals (no reference counterpart)
        const double r = ib_sqrt(nxv * nxv + nyv * nyv);
        if (r > 0) { nxv = nxv / r; nyv = nyv / r; }
    }
    g.nx[o] = nxv;
    g.ny[o] = nyv;
    const int ng = L.ng;
    auto interior = [&](int ii, int jj) { return ii >= ng && ii < L.nx + ng && jj >= ng && jj < L.ny + ng; };
    uint8_t t;
    if (!interior(i, j)) t = 2;
    else if (g.phi[o] >= 0) t = 0;
    else {
        t = 1;
        for (int k = 1; k <= ng; k++) {
            const int ci[4] = {i - k, i + k, i, i};
            const int cj[4] = {j, j, j - k, j + k};
            for (int q = 0; q < 4; q++)
                if (interior(ci[q], cj[q]) && g.phi[L.at(ci[q], cj[q])] >= 0) t = 2;
        }
    }
    g.type[o] = t;
}

// Mirror_IDX.cpp:21-40: the ii in [0, M-1) with x_ii <= xm < x_{ii+1}; the same
// half-open test on the candidates around the closed-form estimate. No match
// leaves 0 (the value a zero-initialised record keeps in the reference).
__device__ int donor_idx(double v, double base, double h, int n) {
    double t = __builtin_floor((v - base) / h);
    int r = 0;
    if (!(t == t)) return 0;
    int t0 = (int)fmax(fmin(t, (double)n + 4.0), -4.0);
    for (int ii = t0 - 2; ii <= t0 + 2; ii++)
        if (ii >= 0 && ii < n - 1 && v < base + (ii + 1) * h && v >= base + ii * h) r = ii;
    return r;
}

// Scheme_Index for one ghost: IDWBaseGFM/GFM_Method.cpp:20-110,
// LinearBaseGFM/GFM_Method.cpp:19-44 + GFM_Bound.cpp:93-575 (donor choice),
// BCGFM/BCGFM_Method.cpp:49-205, HBCGFM/HBCGFM_Method.cpp:56-91.
__global__ void k_geom_ghosts(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const long long o = L.at(i, j);
    if (g.type[o] != 2) return;
    Ghost gh;
    gh.cell = o;
    gh.nx = g.nx[o];
    gh.ny = g.ny[o];
    double ph = ib_fabs(g.phi[o]);
    gh.phi = ph;
    double x = xc(g, i), y = yc(g, j);
    const int bc = g.bc;
    double dxg = xc(g, 1) - xc(g, 0), dyg = yc(g, 2) - yc(g, 1);
    if (bc == IBTFO_SIMPLE_BOUND) {
        // SimpleBound/Scheme_Index.cpp:12-66: the nearest fluid cell within ng
        // along -x (else +x) and -y (else +y); k cells away -> donor offset 1 - 2k
        // (2k - 1). Dependencies on earlier-filled ghosts are resolved on the host.
        auto t = [&](int ii, int jj) { return g.type[L.at(ii, jj)]; };
        int IDX = 0, IDY = 0;
        for (int k = 0; k < ng + 1; k++)
            if (t(i - k, j) < 1) { IDX = 1 - 2 * k; break; }
        if (IDX == 0)
            for (int k = 0; k < ng + 1; k++)
                if (t(i + k, j) < 1) { IDX = 2 * k - 1; break; }
        for (int k = 0; k < ng + 1; k++)
            if (t(i, j - k) < 1) { IDY = 1 - 2 * k; break; }
        if (IDY == 0)
            for (int k = 0; k < ng + 1; k++)
                if (t(i, j + k) < 1) { IDY = 2 * k - 1; break; }
        auto ring_cell = [&](int ii, int jj) { return ii < ng || ii >= L.nx + ng || jj < ng || jj >= L.ny + ng; };
        gh.idx = IDX; gh.idy = IDY;
        gh.d1 = L.at(i + IDX, j);
        gh.d1e = L.at(i, j + IDY);
        gh.idxe = gh.idye = -1;
        gh.pick = 0;
        gh.ring = (ring_cell(i + IDX, j) ? 1 : 0) | (ring_cell(i, j + IDY) ? 16 : 0);
        gh.xm = gh.ym = gh.xe = gh.ye = 0.0;
        for (int q = 0; q < 4; q++) { gh.a[q] = 0.0; gh.ae[q] = 0.0; }
        unsigned int slot = atomicAdd(g.count, 1u);
        if (slot < (unsigned int)g.cap) g.ghosts[slot] = gh;
        return;
    }
    if (bc == 1 || bc == 2 || bc == 3) {
        gh.xm = x + 2 * gh.nx * ph;
        gh.ym = y + 2 * gh.ny * ph;
        gh.xe = gh.ye = 0.0;
    } else {
        double delta = ib_sqrt(dxg * dxg + dyg * dyg) / 2;
        if (bc == 5 && ph > delta) {
            gh.xm = x + gh.nx * (2 * ph); gh.ym = y + gh.ny * (2 * ph);
            gh.xe = x + gh.nx * (2 * ph); gh.ye = y + gh.ny * (2 * ph);
        } else {
            gh.xm = x + gh.nx * (ph + delta); gh.ym = y + gh.ny * (ph + delta);
            gh.xe = x + gh.nx * (ph + 2.0 * delta); gh.ye = y + gh.ny * (ph + 2.0 * delta);
        }
    }
    gh.idx = donor_idx(gh.xm, g.Li, g.dx, L.M);
    gh.idy = donor_idx(gh.ym, g.Hi, g.dx, L.N);
    gh.d1 = L.at(gh.idx, gh.idy);
    const int di[4] = {0, 0, 1, 1}, dj[4] = {0, 1, 1, 0};
    for (int q = 0; q < 4; q++) { gh.a[q] = 0.0; gh.ae[q] = 0.0; }
    gh.idxe = gh.idye = 0;
    gh.d1e = gh.d1;
    if (bc == 1 || bc >= 4) {
        for (int q = 0; q < 4; q++) {
            int ii = gh.idx + di[q], jj = gh.idy + dj[q];
            uint8_t t = g.type[L.at(ii, jj)];
            bool keep = (bc == 1) ? (g.cut ? (t == 0 || t == 3) : (t < 1)) : (t != 1);
            if (keep) {
                double ddx = xc(g, ii) - gh.xm, ddy = yc(g, jj) - gh.ym;
                gh.a[q] = 1 / ib_sqrt(ddx * ddx + ddy * ddy);
            }
        }
    }
    if (bc >= 4) {
        gh.idxe = donor_idx(gh.xe, g.Li, g.dx, L.M);
        gh.idye = donor_idx(gh.ye, g.Hi, g.dx, L.N);
        gh.d1e = L.at(gh.idxe, gh.idye);
        for (int q = 0; q < 4; q++) {
            int ii = gh.idxe + di[q], jj = gh.idye + dj[q];
            if (g.type[L.at(ii, jj)] != 1) {
                double ddx = xc(g, ii) - gh.xe, ddy = yc(g, jj) - gh.ye;
                gh.ae[q] = 1 / ib_sqrt(ddx * ddx + ddy * ddy);
            }
        }
    }
    gh.ring = 0;
    auto in_ring = [&](int ii, int jj) { return ii < ng || ii >= L.nx + ng || jj < ng || jj >= L.ny + ng; };
    for (int q = 0; q < 4; q++) {
        if (in_ring(gh.idx + di[q], gh.idy + dj[q])) gh.ring |= 1 << q;
        if (bc >= 4 && in_ring(gh.idxe + di[q], gh.idye + dj[q])) gh.ring |= 1 << (4 + q);
    }
    gh.pick = -1;
    if (bc == 2 || bc == 3) {
        bool f[4];
        for (int q = 0; q < 4; q++) f[q] = g.type[L.at(gh.idx + di[q], gh.idy + dj[q])] < 1;
        bool f1 = f[0], f2 = f[1], f3 = f[2], f4 = f[3];
        double x_d = gh.xm - xc(g, gh.idx);
        double y_d = gh.xm - xc(g, gh.idx);  // the reference measures y_d with the x mirror too
        bool xr = x_d > dxg / 2, yu = y_d > dyg / 2;
        int pick = -1;
        if (f1 && f4 && f2 && f3) pick = yu ? (xr ? 2 : 1) : (xr ? 3 : 0);
        else if (!f1 && f4 && f2 && f3) pick = yu ? (xr ? 2 : 1) : 3;
        else if (f1 && !f4 && f2 && f3) pick = yu ? (xr ? 2 : 1) : 0;
        else if (f1 && f4 && !f2 && f3) pick = yu ? 2 : (xr ? 3 : 0);
        else if (f1 && f4 && f2 && !f3) pick = yu ? 1 : (xr ? 3 : 0);
        else if (f1 && f4 && !f2 && !f3) pick = xr ? 3 : 0;
        else if (!f1 && !f4 && f2 && f3) pick = xr ? 2 : 1;
        else if (f1 && !f4 && f2 && !f3) pick = yu ? 1 : 0;
        else if (!f1 && f4 && !f2 && f3) pick = yu ? 2 : 3;
        else if (f1 && !f4 && !f2 && f3) pick = yu ? 2 : 0;
        else if (!f1 && f4 && f2 && !f3) pick = yu ? 1 : 3;
        else if (f1 && !f4 && !f2 && !f3) pick = 0;
        else if (!f1 && !f4 && f2 && !f3) pick = 1;
        else if (!f1 && !f4 && !f2 && f3) pick = 2;
        else if (!f1 && f4 && !f2 && !f3) pick = 3;
        gh.pick = pick;
        gh.a[0] = pick;
    }
    unsigned int slot = atomicAdd(g.count, 1u);
    if (slot < (unsigned int)g.cap) g.ghosts[slot] = gh;
}

hipError_t geom_run(GeomArgs& g, hipStream_t s) {
    dim3 blk(128), grd((g.L.M + 127) / 128, g.L.N);
    hipLaunchKernelGGL(k_geom_phi, grd, blk, 0, s, g);
    hipLaunchKernelGGL(k_geom_classify, grd, blk, 0, s, g);
    return hipGetLastError();
}
hipError_t geom_ghosts(GeomArgs& g, hipStream_t s) {
    dim3 blk(128), grd((g.L.M + 127) / 128, g.L.N);
    hipLaunchKernelGGL(k_geom_ghosts, grd, blk, 0, s, g);
    return hipGetLastError();
}



// ------------------------------------------------------------ cut cells
// No reference counterpart. From the cell-centred phi: corner values are the
// means of the four surrounding cells' phi, a face's aperture is its fluid
// fraction under linear interpolation between its corners, the volume fraction
// is the area of the fluid polygon those crossings cut from the cell, and the
// wall vector S = -((aE - aW) dy, (aN - aS) dx) closes the faces.
__device__ __forceinline__ double cut_frac(double p, double q) {
    if (p >= 0 && q >= 0) return 1.0;
    if (p < 0 && q < 0) return 0.0;
    return p >= 0 ? p / (p - q) : q / (q - p);
}
__device__ __forceinline__ double corner_phi(const GeomArgs& g, int i, int j) {  // corner (i+1/2, j+1/2)
    const Layout& L = g.L;
    return 0.25 * (((g.phi[L.at(i, j)] + g.phi[L.at(i + 1, j)]) + g.phi[L.at(i, j + 1)]) + g.phi[L.at(i + 1, j + 1)]);
}
__device__ __forceinline__ long long cell_q(const Layout& L, int i, int j) { return (long long)j * L.M + i; }

__global__ void k_cut_cells(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    if (i >= L.M) return;
    double* cc = g.dcut + cell_q(L, i, j) * 8;
    for (int k = 0; k < 7; k++) cc[k] = 0.0;
    cc[7] = -1.0;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) { g.type[L.at(i, j)] = 2; return; }
    const double sw = corner_phi(g, i - 1, j - 1), se = corner_phi(g, i, j - 1), ne = corner_phi(g, i, j),
                 nw = corner_phi(g, i - 1, j);
    const double aw = cut_frac(sw, nw), ae = cut_frac(se, ne), as = cut_frac(sw, se), an = cut_frac(nw, ne);
    double alpha;
    if (sw >= 0 && se >= 0 && ne >= 0 && nw >= 0) alpha = 1.0;
    else if (sw < 0 && se < 0 && ne < 0 && nw < 0) alpha = 0.0;
    else {
        const double cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1}, cv[4] = {sw, se, ne, nw};
        double px[8], py[8];
        int n = 0;
        for (int e = 0; e < 4; e++) {
            const int f = (e + 1) % 4;
            if (cv[e] >= 0) { px[n] = cx[e]; py[n] = cy[e]; n++; }
            if ((cv[e] >= 0) != (cv[f] >= 0)) {
                const double t = cv[e] / (cv[e] - cv[f]);
                px[n] = cx[e] + t * (cx[f] - cx[e]);
                py[n] = cy[e] + t * (cy[f] - cy[e]);
                n++;
            }
        }
        double a2 = 0.0;
        for (int k = 0; k < n; k++) {
            const int l = (k + 1) % n;
            a2 += px[k] * py[l] - px[l] * py[k];
        }
        alpha = 0.5 * a2;
    }
    cc[0] = alpha; cc[1] = aw; cc[2] = ae; cc[3] = as; cc[4] = an;
    cc[5] = -((ae - aw) * g.cdy);
    cc[6] = -((an - as) * g.cdx);
    g.type[L.at(i, j)] = alpha >= 1.0 ? 0 : (alpha > 0.0 ? 3 : 1);
}

// A cut cell with alpha < 1/2 picks a neighbour with alpha >= 1/2: along the
// normal's dominant axis first, then the other axis, then the opposite
// directions; without one it becomes inactive.
__global__ void k_cut_target(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const long long o = L.at(i, j);
    double* cc = g.dcut + cell_q(L, i, j) * 8;
    if (g.type[o] != 3 || cc[0] >= 0.5) return;
    const double n_x = g.nx[o], n_y = g.ny[o];
    const int sx = n_x >= 0 ? 1 : -1, sy = n_y >= 0 ? 1 : -1;
    int ci[4], cj[4];
    if (ib_fabs(n_x) >= ib_fabs(n_y)) {
        ci[0] = i + sx; cj[0] = j; ci[1] = i; cj[1] = j + sy; ci[2] = i - sx; cj[2] = j; ci[3] = i; cj[3] = j - sy;
    } else {
        ci[0] = i; cj[0] = j + sy; ci[1] = i + sx; cj[1] = j; ci[2] = i; cj[2] = j - sy; ci[3] = i - sx; cj[3] = j;
    }
    for (int q = 0; q < 4; q++) {
        if (ci[q] < ng || ci[q] >= L.nx + ng || cj[q] < ng || cj[q] >= L.ny + ng) continue;
        const uint8_t tq = g.type[L.at(ci[q], cj[q])];
        if ((tq == 0 || tq == 3) && g.dcut[cell_q(L, ci[q], cj[q]) * 8] >= 0.5) {
            cc[7] = (double)cell_q(L, ci[q], cj[q]);
            return;
        }
    }
    g.type[o] = 1;  // no partner: inactive
}

// inactive cells within ng of an active cell along an axis are ghosts
__global__ void k_cut_ghosts(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const long long o = L.at(i, j);
    if (g.type[o] != 1) return;
    for (int k = 1; k <= ng; k++) {
        const int ci[4] = {i - k, i + k, i, i}, cj[4] = {j, j, j - k, j + k};
        for (int q = 0; q < 4; q++) {
            if (ci[q] < ng || ci[q] >= L.nx + ng || cj[q] < ng || cj[q] >= L.ny + ng) continue;
            const uint8_t tq = g.type[L.at(ci[q], cj[q])];
            if (tq == 0 || tq == 3) { g.type[o] = 2; return; }
        }
    }
}

// A face between an active cell and an inactive interior one (zero-area
// polygons where corners are exactly 0, small cells without a partner) is a wall:
// its aperture becomes 0 and the wall vector is closed again, so no flux leaves
// the fluid through it. Each cell writes only its own record; types are read.
__global__ void k_cut_close(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const uint8_t t = g.type[L.at(i, j)];
    if (t != 0 && t != 3) return;
    double* cc = g.dcut + cell_q(L, i, j) * 8;
    const int ni[4] = {i - 1, i + 1, i, i}, nj[4] = {j, j, j - 1, j + 1};
    bool closed = false;
    for (int q = 0; q < 4; q++) {
        if (ni[q] < ng || ni[q] >= L.nx + ng || nj[q] < ng || nj[q] >= L.ny + ng) continue;
        const uint8_t tq = g.type[L.at(ni[q], nj[q])];
        if (tq == 0 || tq == 3 || cc[1 + q] == 0.0) continue;
        cc[1 + q] = 0.0;
        closed = true;
    }
    if (!closed) return;
    cc[5] = -((cc[2] - cc[1]) * g.cdy);
    cc[6] = -((cc[4] - cc[3]) * g.cdx);
}

// a full cell (type 0, all apertures 1) with a face closed above takes the
// cut-cell update: type 3 with alpha = 1
__global__ void k_cut_retype(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const long long o = L.at(i, j);
    const double* cc = g.dcut + cell_q(L, i, j) * 8;
    if (g.type[o] == 0 && (cc[1] == 0.0 || cc[2] == 0.0 || cc[3] == 0.0 || cc[4] == 0.0)) g.type[o] = 3;
}

// compaction: records of the type-3 cells, and merge groups gathered by each
// target from its W, E, S, N neighbours (counting only when the capacity is 0)
__global__ void k_cut_compact(GeomArgs g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const Layout L = g.L;
    const int ng = L.ng;
    if (i < ng || i >= L.nx + ng || j < ng || j >= L.ny + ng) return;
    const long long o = L.at(i, j);
    const uint8_t t = g.type[o];
    const double* cc = g.dcut + cell_q(L, i, j) * 8;
    if (t == 3) {
        const unsigned int slot = atomicAdd(g.ccount, 1u);
        if (slot < (unsigned int)g.ccap) {
            CutCell r;
            r.alpha = cc[0]; r.aw = cc[1]; r.ae = cc[2]; r.as = cc[3]; r.an = cc[4]; r.sbx = cc[5]; r.sby = cc[6];
            r.target = cc[7] < 0 ? -1 : L.at((int)((long long)cc[7] % L.M), (int)((long long)cc[7] / L.M));
            g.cuts[slot] = r;
            g.cut_index[o] = (int)slot;
        }
    }
    if ((t == 0 || t == 3) && cc[0] >= 0.5) {
        const int ni[4] = {i - 1, i + 1, i, i}, nj[4] = {j, j, j - 1, j + 1};
        MergeGroup gr;
        gr.n = 0;
        gr.pad = 0;
        gr.target = o;
        gr.at = cc[0];
        const double me = (double)cell_q(L, i, j);
        for (int q = 0; q < 4; q++) {
            const long long nq = L.at(ni[q], nj[q]);
            const double* cn = g.dcut + cell_q(L, ni[q], nj[q]) * 8;
            if (g.type[nq] == 3 && cn[7] == me) {
                gr.member[gr.n] = nq;
                gr.am[gr.n] = cn[0];
                gr.n++;
            }
        }
        for (int q = gr.n; q < 4; q++) { gr.member[q] = o; gr.am[q] = 0.0; }
        if (gr.n > 0) {
            const unsigned int slot = atomicAdd(g.gcount, 1u);
            if (slot < (unsigned int)g.gcap) g.groups[slot] = gr;
        }
    }
}

hipError_t cut_run(GeomArgs& g, hipStream_t s) {
    dim3 blk(128), grd((g.L.M + 127) / 128, g.L.N);
    hipLaunchKernelGGL(k_cut_cells, grd, blk, 0, s, g);
    hipLaunchKernelGGL(k_cut_target, grd, blk, 0, s, g);
    hipLaunchKernelGGL(k_cut_ghosts, grd, blk, 0, s, g);
    hipLaunchKernelGGL(k_cut_close, grd, blk, 0, s, g);
    hipLaunchKernelGGL(k_cut_retype, grd, blk, 0, s, g);
    return hipGetLastError();
}
hipError_t cut_compact(GeomArgs& g, hipStream_t s) {
    dim3 blk(128), grd((g.L.M + 127) / 128, g.L.N);
    hipLaunchKernelGGL(k_cut_compact, grd, blk, 0, s, g);
    return hipGetLastError();
}

// ------------------------------------------------------------ mask level set
// Test/PNGGeo/Level_Set.cpp:31-149 on the device. The reference computes, for
// every cell, the distance to every edge pixel (O(cells x edges)); here the
// nearest edge pixel is found by a separable exact search (per pixel column the
// nearest edge row to each cell row, then per cell row the lower envelope of the
// parabolas (x_q - x_e)^2 + g(x_e) over pixel columns, Felzenszwalb and
// Huttenlocher, "Distance Transforms of Sampled Functions", 2012), and the
// distance to that pixel is evaluated with the reference's expression.

// Level_Set.cpp:36-48: +1 where R + G + B > 0, else -1; image rows flipped so that
// row 0 is the bottom of the domain
__global__ void k_mask_sign(MaskArgs m) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= m.W) return;
    const unsigned char* px = m.rgba + ((long long)y * m.W + x) * 4;
    const double s = (double)px[0] + (double)px[1] + (double)px[2];
    m.v[(long long)(m.H - y - 1) * m.W + x] = (s > 0) * 2 - 1;
}

// Level_Set.cpp:51-76: edge pixel where max(|v - v_left|, |v - v_below|) > 0.1,
// pixels 1..W-2 x 1..H-2 only
__global__ void k_mask_edge(MaskArgs m) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= m.W) return;
    unsigned char e = 0;
    if (x >= 1 && x < m.W - 1 && y >= 1 && y < m.H - 1) {
        const double c = m.v[(long long)y * m.W + x];
        const double dx = c - m.v[(long long)y * m.W + x - 1];
        const double dy = c - m.v[(long long)(y - 1) * m.W + x];
        const double g = (ib_fabs(dx) > ib_fabs(dy)) ? ib_fabs(dx) : ib_fabs(dy);
        e = g > 0.1;
    }
    m.edge[(long long)y * m.W + x] = e;
}

__device__ __forceinline__ double edge_x(const MaskArgs& m, int x) { return m.exo + m.dxp * x; }
__device__ __forceinline__ double edge_y(const MaskArgs& m, int y) { return m.eyo + m.dyp * y; }

// Per pixel column: the first and last edge row of each 32-row segment.
__global__ void k_edt_segs(MaskArgs m) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, sg = blockIdx.y;
    if (x >= m.W) return;
    int first = -1, last = -1;
    const int y1 = min((sg + 1) * MASK_SEG, m.H);
    for (int y = sg * MASK_SEG; y < y1; y++)
        if (m.edge[(long long)y * m.W + x]) {
            if (first < 0) first = y;
            last = y;
        }
    m.sfirst[(long long)sg * m.W + x] = first;
    m.slast[(long long)sg * m.W + x] = last;
}

// Per pixel column: the last edge row below each segment and the first above it.
__global__ void k_edt_segscan(MaskArgs m) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m.W) return;
    int b = -1;
    for (int sg = 0; sg < m.nseg; sg++) {
        const long long o = (long long)sg * m.W + x;
        m.before[o] = b;
        if (m.slast[o] >= 0) b = m.slast[o];
    }
    int a = -1;
    for (int sg = m.nseg - 1; sg >= 0; sg--) {
        const long long o = (long long)sg * m.W + x;
        m.after[o] = a;
        if (m.sfirst[o] >= 0) a = m.sfirst[o];
    }
}

// Per (pixel column x, cell row j): the column's edge rows just below and just
// above the cell-centre height y_j (the last with edge height <= y_j and the
// first above it: within the 32-row segment around y_j, else from the segment
// summaries), the nearer of them and its squared height difference.
__global__ void k_edt_cols(MaskArgs m) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)m.W * m.N) return;
    const int x = (int)(t % m.W), j = (int)(t / m.W);
    const double yq = m.Hi + j * m.dxc;
    // y0: the last pixel row whose edge height is <= yq (-1: none); the heights
    // increase with the row, so a rounded estimate is corrected by single steps
    int y0 = (int)floor((yq - m.eyo) / m.dyp);
    y0 = min(max(y0, -1), m.H - 1);
    while (y0 >= 0 && edge_y(m, y0) > yq) y0--;
    while (y0 + 1 < m.H && edge_y(m, y0 + 1) <= yq) y0++;
    int prv = -1, nxt = -1;
    if (y0 >= 0) {
        const int s0 = y0 / MASK_SEG;
        for (int y = y0; y >= s0 * MASK_SEG; y--)
            if (m.edge[(long long)y * m.W + x]) { prv = y; break; }
        if (prv < 0) prv = m.before[(long long)s0 * m.W + x];
    }
    if (y0 + 1 < m.H) {
        const int y1 = y0 + 1, s1 = y1 / MASK_SEG, ye = min((s1 + 1) * MASK_SEG, m.H);
        for (int y = y1; y < ye; y++)
            if (m.edge[(long long)y * m.W + x]) { nxt = y; break; }
        if (nxt < 0) nxt = m.after[(long long)s1 * m.W + x];
    }
    double best = HUGE_VAL;
    int arg = -1;
    if (prv >= 0) {
        double d = edge_y(m, prv) - yq;
        best = d * d;
        arg = prv;
    }
    if (nxt >= 0) {
        double d = edge_y(m, nxt) - yq;
        if (d * d < best) { best = d * d; arg = nxt; }
    }
    m.g1[t] = best;
    m.r1[t] = arg;
    m.r2[t] = arg == prv ? nxt : prv;  // the other neighbour (near ties)
}

// phi of cell (i, j) from the lower envelope (v[0..k], boundaries z) of cell row
// j: the envelope's parabola at the cell centre and its neighbours, each with the
// column's two nearest edge rows; points whose distances tie with the nearest in
// exact arithmetic but round differently are all evaluated with the reference's
// expression, whose minimum it keeps. Then the sign sampled at pixel
// (int((i+0.5) W / M), int((j+0.5) H / N)).
template <typename V>
__device__ __forceinline__ double mask_phi_cell(const MaskArgs& m, int i, int j, const V* v, const double* z,
                                                int k) {
    const double X = m.Li + i * m.dxc, Y = m.Hi + j * m.dxc;
    double dmin = 100.0;  // Level_Set.cpp:131 starts every search at 100
    if (k >= 0) {
        // q = the last envelope piece whose left boundary lies below X
        int lo = 0, hi = k;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (z[mid] < X) lo = mid;
            else hi = mid - 1;
        }
        const int q = lo;
        for (int c = (q > 0 ? q - 1 : q); c <= (q < k ? q + 1 : q); c++) {
            const int ex = v[c];
            const double ax = edge_x(m, ex);
            for (int rr = 0; rr < 2; rr++) {
                const int ey = rr == 0 ? m.r1[(long long)j * m.W + ex] : m.r2[(long long)j * m.W + ex];
                if (ey < 0) continue;
                const double ay = edge_y(m, ey);
                const double d = ib_sqrt((ax - X) * (ax - X) + (ay - Y) * (ay - Y));
                if (d < dmin) dmin = d;
            }
        }
    }
    const int xs = (int)((i + 0.5) * (m.W) / (m.M));
    const int ys = (int)((j + 0.5) * (m.H) / (m.N));
    const int pn = m.v[(long long)ys * m.W + xs] >= 0 ? 1 : -1;
    return pn * dmin;
}

// Per cell row j (one 64-lane block): the lower envelope of f_x(q) = (q - x_e(x))^2
// + g1(x, j) over the pixel columns that hold an edge (Felzenszwalb-Huttenlocher),
// built by lane 0 with its stack in LDS, the row's g1 staged in LDS by all lanes;
// then every lane evaluates cells i = lane, lane + 64, ...
__global__ __launch_bounds__(64) void k_edt_rows(MaskArgs m) {
    extern __shared__ double lds[];
    const int j = blockIdx.x, lane = threadIdx.x, W = m.W;
    double* z = lds;                                    // W + 1
    double* gr = z + (W + 1);                           // W
    unsigned short* v = (unsigned short*)(gr + W);      // W + 1
    __shared__ int kk;
    for (int x = lane; x < W; x += 64) gr[x] = m.g1[(long long)j * W + x];
    __syncthreads();
    if (lane == 0) {
        int k = -1;
        for (int x = 0; x < W; x++) {
            const double f = gr[x];
            if (!(f < HUGE_VAL)) continue;
            const double p = edge_x(m, x);
            double s = -HUGE_VAL;
            while (k >= 0) {
                const int u = v[k];
                const double pu = edge_x(m, u), fu = gr[u];
                s = ((f + p * p) - (fu + pu * pu)) / (2.0 * (p - pu));
                if (s <= z[k]) k--;
                else break;
            }
            k++;
            v[k] = (unsigned short)x;
            z[k] = k == 0 ? -HUGE_VAL : s;
        }
        kk = k;
    }
    __syncthreads();
    const int k = kk;
    for (int i = lane; i < m.M; i += 64) m.phi[(long long)j * m.M + i] = mask_phi_cell(m, i, j, v, z, k);
}

// The same for masks too wide for the LDS envelope: one thread per cell row,
// its stack in global scratch.
__global__ void k_edt_rows_global(MaskArgs m) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.N) return;
    int* v = m.vbuf + (long long)j * (m.W + 1);
    double* z = m.zbuf + (long long)j * (m.W + 1);
    const double* g = m.g1 + (long long)j * m.W;
    int k = -1;
    for (int x = 0; x < m.W; x++) {
        const double f = g[x];
        if (!(f < HUGE_VAL)) continue;
        const double p = edge_x(m, x);
        double s = -HUGE_VAL;
        while (k >= 0) {
            const int u = v[k];
            const double pu = edge_x(m, u), fu = g[u];
            s = ((f + p * p) - (fu + pu * pu)) / (2.0 * (p - pu));
            if (s <= z[k]) k--;
            else break;
        }
        k++;
        v[k] = x;
        z[k] = k == 0 ? -HUGE_VAL : s;
    }
    for (int i = 0; i < m.M; i++) m.phi[(long long)j * m.M + i] = mask_phi_cell(m, i, j, v, z, k);
}

hipError_t mask_run(MaskArgs& m, hipStream_t s) {
    dim3 pg((m.W + 255) / 256, m.H);
    hipLaunchKernelGGL(k_mask_sign, pg, dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_mask_edge, pg, dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_edt_segs, dim3((m.W + 255) / 256, m.nseg), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_edt_segscan, dim3((m.W + 255) / 256), dim3(256), 0, s, m);
    const long long nq = (long long)m.W * m.N;
    hipLaunchKernelGGL(k_edt_cols, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, m);
    if (!m.vbuf) {
        hipLaunchKernelGGL(k_edt_rows, dim3(m.N), dim3(64), mask_lds_bytes(m.W), s, m);
    } else {
        hipLaunchKernelGGL(k_edt_rows_global, dim3((m.N + 63) / 64), dim3(64), 0, s, m);
    }
    return hipGetLastError();
}
const void* mask_rows_kernel() { return (const void*)k_edt_rows; }

}  // namespace ibtfo
