// IBTFO CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's per-timestep Euler finite-volume
// update, used by tests/, by __graft_entry__.smoke() and by bench.py's CPU
// baseline leg as the result to compare against. Nothing in the product
// (ibtfo_amd/, include/) links, loads or calls this file.
//
// Layout follows the reference exactly: AoS state `U[(M*j+i)*4+k]`, geometry
// `XY[(M*j+i)*6+c]` with c = {x, y, phi, n_x, n_y, type}, a scratch record
// `T[(M*j+i)*W + slot*4 + k]` whose slot pairs alias the way the reference's
// EQDefine.H headers do (F_OLD==G_OLD, F_L==G_D, F_R==G_U, U_L==U_D, U_R==U_U),
// and the same loop ranges and arithmetic order, so results are bit-identical
// with the reference built by g++ (-ffp-contract=off) on the same inputs.
// Each function cites the reference file:line it restates.
//
// Pinning (what checks each function; see tests/):
//   parse/limiters/fluxes/solvers/schemes/dt/BCs/geometry: bit-exact against
//   the reference translation units compiled from source (oracle/_ref) and the
//   committed golden fixtures (tests/golden/); plus closed forms (textbook
//   Euler flux, HLL consistency F(U,U)=F(U), Steger-Warming F+ + F- = F,
//   free-stream preservation, exact Sod solution trend).
//   PVRS Y stray store (PVRS/Riemann_solver.cpp:550) is restated with the
//   sequential (g++ host) ordering; the reference's parallel builds race there.
//   Functions with no reference counterpart (periodic / inflow edge policies,
//   isentropic-vortex and x-split Sod initialisers, cut-cell): no reference
//   output exists; each is pinned by closed forms and invariants instead
//   (tests/test_oracle_pins.py, tests/test_oracle_properties.py,
//   tests/test_oracle_cutcell.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { SCH_LAX = 0, SCH_MUSCL = 1, SCH_MHM = 2, SCH_GODUNOV = 3 };
enum { RS_HLL = 0, RS_HLLC = 1, RS_PVRS = 2, RS_SW = 3 };
enum { LIM_MINMOD = 0, LIM_VANLEER = 1, LIM_VANALBADA = 2 };
enum { BC_REGULAR = 0, BC_IDW = 1, BC_LINEAR = 2, BC_RIEMANN = 3, BC_BCGFM = 4, BC_HBCGFM = 5, BC_CUT = 6,
       BC_SIMPLE = 7 };
enum { EDGE_REFERENCE = 0, EDGE_PERIODIC = 1, EDGE_INFLOW = 2, EDGE_PNGGEO = 3 };

// Sound speed. The reference computes std::pow(x, 0.5) (HLL/Riemann_solver.cpp:144,
// ComputeDt.cpp:42); glibc's pow is not always the correctly rounded root. Mode 1
// swaps in sqrt so a run can be compared bit-for-bit with an IEEE sqrt device path.
int g_root_mode = 0;
inline double root(double x) { return g_root_mode ? std::sqrt(x) : std::pow(x, 0.5); }

// cut-cell records of the current solve (or_set_cut), 8 doubles per cell:
// alpha, aperture W, E, S, N, wall vector x, y, merge target (-1: none)
const double* g_cut = nullptr;

struct Grid {
    int nx, ny, ng, M, N;
    Grid(int nx_, int ny_, int ng_) : nx(nx_), ny(ny_), ng(ng_), M(nx_ + 2 * ng_), N(ny_ + 2 * ng_) {}
    size_t c(int i, int j) const { return (size_t)M * j + i; }
};

// ---------------------------------------------------------------- limiters
// minmod/Slope_limiter.H:4
inline double lim_minmod(double a, double b) { return (a * b > 0) ? ((a / b > 1) ? b : a) : 0; }
// VanLeer/Slope_limiter.H:4 as the preprocessor expands it at every call site: the
// second argument is always written `P - Q` without parentheses, so the macro's
// `* b /` becomes `* P - Q /`.
inline double lim_vanleer_ref(double a, double P, double Q) {
    double r = (a + 1e-7) / ((P - Q) + 1e-7);
    return (r > 0) ? ((r * r + r) * P - Q / (r * r + 1)) : 0;
}
// The formula VanLeer/Slope_limiter.H:4 writes, with its argument parenthesised
// (corrected mode), in one-division form: with A = a + 1e-7, B = b + 1e-7 and
// r = A/B, (r*r + r) b / (r*r + 1) = b A (A + B) / (A*A + B*B) and r > 0 iff
// A*B > 0. The two forms agree to a few ulp (tests pin the identity); this one
// is the corrected mode's definition.
inline double lim_vanleer_fixed(double a, double b) {
    double A = a + 1e-7, B = b + 1e-7;
    return (A * B > 0) ? (b * (A * (A + B)) / (A * A + B * B)) : 0;
}
// VanAlbada/Slope_limiter.H:4 does not compile (undeclared r); this is its evident
// intent with r defined as in the sibling limiter. No reference output exists.
inline double lim_vanalbada(double a, double b) {
    double r = (a + 1e-7) / (b + 1e-7);
    return (r > 0) ? ((r + std::fabs(r)) * b / (1.0 + std::fabs(r))) : 0;
}
// slope for the difference pair (a, b = P - Q)
inline double limit(int lim, int quirks, double a, double P, double Q) {
    switch (lim) {
        case LIM_MINMOD: return lim_minmod(a, P - Q);
        case LIM_VANLEER: return quirks ? lim_vanleer_ref(a, P, Q) : lim_vanleer_fixed(a, P - Q);
        default: return lim_vanalbada(a, P - Q);
    }
}

struct Ctx {
    Grid g;
    double L, H, gam, dx, dy;
    int W;            // doubles per scratch record: 28 (MUSCL/MHM/Godunov) or 16 (Lax)
    double* T;        // scratch record array
    int quirks;
    Ctx(int nx, int ny, int ng, double L_, double H_, double gam_, int W_, double* T_, int q)
        : g(nx, ny, ng), L(L_), H(H_), gam(gam_), dx(L_ / nx), dy(H_ / ny), W(W_), T(T_), quirks(q) {}
    double* slot(int i, int j, int s) const { return T + g.c(i, j) * W + 4 * s; }
};

enum { S_TMP = 0, S_TMPRK = 1, S_FOLD = 2, S_GOLD = 2, S_FL = 3, S_FR = 4, S_GD = 3, S_GU = 4,
       S_UL = 5, S_UR = 6, S_UD = 5, S_UU = 6 };

// ------------------------------------------------------------ physical flux
// Conserve2Flux.cpp:59-76 (x) and :94-111 (y)
inline void flux_x(double gam, const double* U, double* F) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho;
    double p = (gam - 1) * (U[3] - 0.5 * rho * (u * u + v * v));
    F[0] = rho * u; F[1] = rho * u * u + p; F[2] = rho * u * v; F[3] = (U[3] + p) * u;
}
inline void flux_y(double gam, const double* U, double* G) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho;
    double p = (gam - 1) * (U[3] - 0.5 * rho * (u * u + v * v));
    G[0] = rho * v; G[1] = rho * u * v; G[2] = rho * v * v + p; G[3] = (U[3] + p) * v;
}
void conserve2flux_x(Ctx& c, int src, int dst) {
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) flux_x(c.gam, c.slot(i, j, src), c.slot(i, j, dst));
}
void conserve2flux_y(Ctx& c, int src, int dst) {
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) flux_y(c.gam, c.slot(i, j, src), c.slot(i, j, dst));
}

// primitive state of a conserved 4-vector with the normal velocity first
struct Prim { double rho, un, ut, p; };
inline Prim prim_x(double gam, const double* U) {
    Prim s; s.rho = U[0]; s.un = U[1] / s.rho; s.ut = U[2] / s.rho;
    s.p = (gam - 1) * (U[3] - 0.5 * s.rho * (s.un * s.un + s.ut * s.ut)); return s;
}
inline Prim prim_y(double gam, const double* U) {
    Prim s; s.rho = U[0]; double u = U[1] / s.rho, v = U[2] / s.rho;
    s.un = v; s.ut = u; s.p = (gam - 1) * (U[3] - 0.5 * s.rho * (u * u + v * v)); return s;
}

// ------------------------------------------------------------------ HLL
// HLL/Riemann_solver.cpp:126-172 (x) and :196-238 (y). `lo` is the left/lower state.
// Returns 1 when the first branch (S_lo >= 0) is taken.
// quirks = 0: the corrected estimate (Toro, Riemann Solvers, 3rd ed., eq. 10.59-10.60):
// p* clamped at 0 and the shock factor applied only where p* > p_K. No reference output.
inline int hll_speeds(double gam, const Prim& a, const Prim& b, double& S_lo, double& S_hi, int quirks = 1) {
    double a_lo = root(gam * a.p / a.rho), a_hi = root(gam * b.p / b.rho);
    double rho_bar = 0.5 * (a.rho + b.rho);
    double a_bar = 0.5 * (a_lo + a_hi);
    double p_star = 0.5 * (a.p + b.p) + 0.5 * (a.un - b.un) * (rho_bar * a_bar);
    if (quirks) {
        S_lo = a.un - a_lo * std::sqrt((gam + 1) * p_star / 2 / gam / a.p + (gam - 1) / 2 / gam);
        S_hi = b.un + a_hi * std::sqrt((gam + 1) * p_star / 2 / gam / b.p + (gam - 1) / 2 / gam);
    } else {
        p_star = p_star > 0.0 ? p_star : 0.0;
        double q_lo = p_star <= a.p ? 1.0 : std::sqrt((gam + 1) * p_star / 2 / gam / a.p + (gam - 1) / 2 / gam);
        double q_hi = p_star <= b.p ? 1.0 : std::sqrt((gam + 1) * p_star / 2 / gam / b.p + (gam - 1) / 2 / gam);
        S_lo = a.un - a_lo * q_lo;
        S_hi = b.un + a_hi * q_hi;
    }
    return S_lo >= 0;
}
void riemann_hll_x(Ctx& c) {
    conserve2flux_x(c, S_UL, S_FL);
    conserve2flux_x(c, S_UR, S_FR);
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            const double *UL = c.slot(i, j, S_UL), *UR = c.slot(i, j, S_UR);
            const double *FL = c.slot(i, j, S_FL), *FR = c.slot(i, j, S_FR);
            double* F = c.slot(i, j, S_FOLD);
            double SL, SR;
            hll_speeds(c.gam, prim_x(c.gam, UL), prim_x(c.gam, UR), SL, SR, c.quirks);
            for (int k = 0; k < 4; k++) {
                if (SL >= 0) F[k] = FL[k];
                else if (SR <= 0) F[k] = FR[k];
                else F[k] = (SR * FL[k] - SL * FR[k] + SR * SL * (UR[k] - UL[k])) / (SR - SL);
            }
        }
}
void riemann_hll_y(Ctx& c) {
    conserve2flux_y(c, S_UD, S_GD);
    conserve2flux_y(c, S_UU, S_GU);
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            const double *UD = c.slot(i, j, S_UD), *UU = c.slot(i, j, S_UU);
            const double *GD = c.slot(i, j, S_GD), *GU = c.slot(i, j, S_GU);
            double* G = c.slot(i, j, S_GOLD);
            double SD, SU;
            hll_speeds(c.gam, prim_y(c.gam, UD), prim_y(c.gam, UU), SD, SU, c.quirks);
            for (int k = 0; k < 4; k++) {
                if (SD >= 0) {
                    // reference: slot left as is, i.e. the x-flux of this cell (aliased
                    // slot, HLL/Riemann_solver.cpp:224-226); corrected: upwind flux
                    if (!c.quirks) G[k] = GD[k];
                } else if (SU <= 0) G[k] = GU[k];
                else G[k] = (SU * GD[k] - SD * GU[k] + SU * SD * (UU[k] - UD[k])) / (SU - SD);
            }
        }
}

// ------------------------------------------------------------------ PVRS
// PVRS/Riemann_solver.cpp:302-437 (x), :457-591 (y). Output is written through
// put(k, value); the y solver maps the momentum components.
struct PvrsOut { double f[4]; int stray; };
inline void pvrs_face(double gam, const double* Ulo, const double* Uhi, const Prim& a, const Prim& b,
                      int ydir, PvrsOut& o) {
    double a_lo = root(gam * a.p / a.rho), a_hi = root(gam * b.p / b.rho);
    double rho_bar = 0.5 * (a.rho + b.rho);
    double a_bar = 0.5 * (a_lo + a_hi);
    double p_star = 0.5 * (a.p + b.p) + 0.5 * (a.un - b.un) * (rho_bar * a_bar);
    double u_star = 0.5 * (a.un + b.un) + 0.5 * (a.p - b.p) / (rho_bar * a_bar);
    double rlo_star = a.rho + (p_star - a.p) / a_lo / a_lo;
    double rhi_star = b.rho + (p_star - b.p) / a_hi / a_hi;
    double S_lo = a.un - a_lo * std::sqrt((gam + 1) * p_star / 2 / gam / a.p + (gam - 1) / 2 / gam);
    double S_hi = b.un + a_hi * std::sqrt((gam + 1) * p_star / 2 / gam / b.p + (gam - 1) / 2 / gam);
    double S_Hlo = a.un - a_lo;
    double S_Tlo = u_star - std::pow(a_lo * (p_star / b.p), ((gam - 1) / 2 / gam));
    double S_Hhi = b.un + a_hi;
    double S_Thi = u_star + std::pow(a_hi * (p_star / b.p), ((gam - 1) / 2 / gam));
    o.stray = 0;
    // physical flux of one side: normal mass, normal momentum, tangential momentum, energy.
    // The tangential momentum is rho*u*v in both directions (left-to-right products).
    auto side = [&](const Prim& s, const double* U) {
        o.f[0] = s.rho * s.un;
        o.f[1] = s.rho * s.un * s.un + s.p;
        o.f[2] = ydir ? s.rho * s.ut * s.un : s.rho * s.un * s.ut;
        o.f[3] = (U[3] + s.p) * s.un;
    };
    auto star = [&](double rs, double ut) {
        o.f[0] = rs * u_star;
        o.f[1] = rs * u_star * u_star + p_star;
        o.f[2] = rs * u_star * ut;
        o.f[3] = (gam * p_star / (gam - 1) + 0.5 * rs * (u_star * u_star + ut * ut)) * u_star;
    };
    auto fan = [&](const Prim& s, double as, int sgn) {
        double base = 2 / (gam + 1) + (gam - 1) * (s.rho) / (gam + 1) / as;
        double rf = s.rho * std::pow(base, 2 / (gam - 1));
        double uf = 2 * (as + (gam - 1) * s.un / 2) / (gam + 1);
        double pf = s.p * std::pow(base, 2 * gam / (gam - 1));
        (void)sgn;
        o.f[0] = rf * uf;
        o.f[1] = rf * uf * uf + pf;
        o.f[2] = rf * uf * s.ut;
        o.f[3] = (gam * pf / (gam - 1) + 0.5 * rf * (uf * uf + s.ut * s.ut)) * uf;
    };
    if (u_star > 0) {
        if (p_star > a.p) { if (S_lo > 0) side(a, Ulo); else star(rlo_star, a.ut); }
        else {
            if (S_Hlo > 0) side(a, Ulo);
            else if (S_Tlo < 0) star(rlo_star, a.ut);
            else fan(a, a_lo, -1);
        }
    } else {
        if (p_star > b.p) {
            if (S_hi < 0) { side(b, Uhi); o.stray = 1; }
            else star(rhi_star, b.ut);
        } else {
            if (S_Hhi < 0) side(b, Uhi);
            else if (S_Thi > 0) star(rhi_star, b.ut);
            else fan(b, a_hi, 1);
        }
    }
}
void riemann_pvrs_x(Ctx& c) {
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            const double *UL = c.slot(i, j, S_UL), *UR = c.slot(i, j, S_UR);
            PvrsOut o;
            pvrs_face(c.gam, UL, UR, prim_x(c.gam, UL), prim_x(c.gam, UR), 0, o);
            double* F = c.slot(i, j, S_FOLD);
            F[0] = o.f[0]; F[1] = o.f[1]; F[2] = o.f[2]; F[3] = o.f[3];
        }
}
// 1: the reference's stray store of PVRS/Riemann_solver.cpp:550 (reference mode);
// 0: dropped -- the definition the GPU path implements (see riemann_pvrs_y)
int g_pvrs_stray = 1;
long long g_stray_faces = 0, g_stray_gold = 0;  // faces taking the branch; of them, stores landing in a G_OLD slot
void riemann_pvrs_y(Ctx& c) {
    // sequential: the reference's stray store (:550) writes the energy flux of
    // face record (i, j) through the state index macro, Index(i, j, 3) = 4(Mj+i)+3,
    // into the scratch record of cell (Mj+i)/7, slot (Mj+i) % 7; the face's own
    // energy slot keeps the aliased x flux. With g_pvrs_stray = 0 only the latter.
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            const double *UD = c.slot(i, j, S_UD), *UU = c.slot(i, j, S_UU);
            PvrsOut o;
            pvrs_face(c.gam, UD, UU, prim_y(c.gam, UD), prim_y(c.gam, UU), 1, o);
            double* G = c.slot(i, j, S_GOLD);
            // y flux order: mass, x-momentum (tangential), y-momentum (normal), energy
            G[0] = o.f[0]; G[1] = o.f[2]; G[2] = o.f[1];
            if (o.stray && c.quirks) {
                g_stray_faces++;
                g_stray_gold += (c.g.c(i, j) * 4 + 3) % c.W / 4 == S_GOLD;
                if (g_pvrs_stray) c.T[c.g.c(i, j) * 4 + 3] = o.f[3];
            } else {
                G[3] = o.f[3];
            }
        }
}

// --------------------------------------------------------- Steger-Warming
// StegerWarming/Riemann_solver.cpp:173-234 (x), :262-325 (y)
inline void sw_split(double gam, const Prim& s, double h, double a, int plus, double out[4], int ydir) {
    const double delta = 1e-6;
    double l1 = plus ? 0.5 * (s.un + std::sqrt(s.un * s.un + delta * delta))
                     : 0.5 * (s.un - std::sqrt(s.un * s.un + delta * delta));
    double l3 = plus ? 0.5 * ((s.un - a) + std::sqrt((s.un - a) * (s.un - a) + delta * delta))
                     : 0.5 * ((s.un - a) - std::sqrt((s.un - a) * (s.un - a) + delta * delta));
    double l4 = plus ? 0.5 * ((s.un + a) + std::sqrt((s.un + a) * (s.un + a) + delta * delta))
                     : 0.5 * ((s.un + a) - std::sqrt((s.un + a) * (s.un + a) + delta * delta));
    double u2 = ydir ? (s.ut * s.ut + s.un * s.un) : (s.un * s.un + s.ut * s.ut);
    out[0] = s.rho * (2 * (gam - 1) * l1 + l3 + l4) / 2 / gam;
    out[1] = s.rho * (2 * (gam - 1) * s.un * l1 + (s.un - a) * l3 + (s.un + a) * l4) / 2 / gam;
    out[2] = s.rho * (2 * (gam - 1) * s.ut * l1 + l3 * s.ut + l4 * s.ut) / 2 / gam;
    out[3] = s.rho * ((gam - 1) * u2 * l1 + (h - a * s.un) * l3 + (h + a * s.un) * l4) / 2 / gam;
}
inline void sw_face(double gam, const Prim& a, const Prim& b, int ydir, double f[4]) {
    double a_lo = root(gam * a.p / a.rho), a_hi = root(gam * b.p / b.rho);
    double e_lo = a.p / (gam - 1) / a.rho, e_hi = b.p / (gam - 1) / b.rho;
    double ux_lo = ydir ? a.ut : a.un, uy_lo = ydir ? a.un : a.ut;
    double ux_hi = ydir ? b.ut : b.un, uy_hi = ydir ? b.un : b.ut;
    double h_lo = e_lo + ux_lo * ux_lo / 2 + uy_lo * uy_lo / 2 + a.p / a.rho;
    double h_hi = e_hi + ux_hi * ux_hi / 2 + uy_hi * uy_hi / 2 + b.p / b.rho;
    double fp[4], fn[4];
    sw_split(gam, a, h_lo, a_lo, 1, fp, ydir);
    sw_split(gam, b, h_hi, a_hi, 0, fn, ydir);
    for (int k = 0; k < 4; k++) f[k] = fp[k] + fn[k];
}
void riemann_sw_x(Ctx& c) {
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            double f[4];
            sw_face(c.gam, prim_x(c.gam, c.slot(i, j, S_UL)), prim_x(c.gam, c.slot(i, j, S_UR)), 0, f);
            double* F = c.slot(i, j, S_FOLD);
            for (int k = 0; k < 4; k++) F[k] = f[k];
        }
}
void riemann_sw_y(Ctx& c) {
#pragma omp parallel for
    for (int j = 0; j < c.g.N; j++)
        for (int i = 0; i < c.g.M; i++) {
            double f[4];
            sw_face(c.gam, prim_y(c.gam, c.slot(i, j, S_UD)), prim_y(c.gam, c.slot(i, j, S_UU)), 1, f);
            double* G = c.slot(i, j, S_GOLD);
            G[0] = f[0]; G[1] = f[2]; G[2] = f[1]; G[3] = f[3];
        }
}

void riemann_x(Ctx& c, int rs) {
    if (rs == RS_PVRS) riemann_pvrs_x(c);
    else if (rs == RS_SW) riemann_sw_x(c);
    else riemann_hll_x(c);
}
void riemann_y(Ctx& c, int rs) {
    if (rs == RS_PVRS) riemann_pvrs_y(c);
    else if (rs == RS_SW) riemann_sw_y(c);
    else riemann_hll_y(c);
}

// ---------------------------------------------------------------- schemes
// piecewise-linear face states, MUSCL_Scheme.cpp:74-104 / :135-165
void recon_x(Ctx& c, const double* U, int lim, int i_lo_L) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = 0; j < g.N; j++) {
        for (int i = i_lo_L; i < g.M - 1; i++)
            for (int k = 0; k < 4; k++) {
                double u0 = U[g.c(i - 1, j) * 4 + k], u1 = U[g.c(i, j) * 4 + k], u2 = U[g.c(i + 1, j) * 4 + k];
                c.slot(i, j, S_UL)[k] = u1 + 0.5 * limit(lim, c.quirks, u1 - u0, u2, u1);
            }
        for (int i = 0; i < g.M - 2; i++)
            for (int k = 0; k < 4; k++) {
                double u1 = U[g.c(i, j) * 4 + k], u2 = U[g.c(i + 1, j) * 4 + k], u3 = U[g.c(i + 2, j) * 4 + k];
                c.slot(i, j, S_UR)[k] = u2 - 0.5 * limit(lim, c.quirks, u3 - u2, u2, u1);
            }
    }
}
void recon_y(Ctx& c, const double* U, int lim) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = g.ng - 1; j < g.N - 1; j++)
        for (int i = 0; i < g.M; i++)
            for (int k = 0; k < 4; k++) {
                double u0 = U[g.c(i, j - 1) * 4 + k], u1 = U[g.c(i, j) * 4 + k], u2 = U[g.c(i, j + 1) * 4 + k];
                c.slot(i, j, S_UD)[k] = u1 + 0.5 * limit(lim, c.quirks, u1 - u0, u2, u1);
            }
#pragma omp parallel for
    for (int j = 0; j < g.N - 2; j++)
        for (int i = 0; i < g.M; i++)
            for (int k = 0; k < 4; k++) {
                double u1 = U[g.c(i, j) * 4 + k], u2 = U[g.c(i, j + 1) * 4 + k], u3 = U[g.c(i, j + 2) * 4 + k];
                c.slot(i, j, S_UU)[k] = u2 - 0.5 * limit(lim, c.quirks, u3 - u2, u2, u1);
            }
}
// Godunov_Scheme.cpp:30-56 / :85-111
void copy_x(Ctx& c, const double* U) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = 0; j < g.N; j++) {
        for (int i = 1; i < g.M - 1; i++)
            for (int k = 0; k < 4; k++) c.slot(i, j, S_UL)[k] = U[g.c(i, j) * 4 + k];
        for (int i = 0; i < g.M - 2; i++)
            for (int k = 0; k < 4; k++) c.slot(i, j, S_UR)[k] = U[g.c(i + 1, j) * 4 + k];
    }
}
void copy_y(Ctx& c, const double* U) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = g.ng - 1; j < g.N - 1; j++)
        for (int i = 0; i < g.M; i++)
            for (int k = 0; k < 4; k++) c.slot(i, j, S_UD)[k] = U[g.c(i, j) * 4 + k];
#pragma omp parallel for
    for (int j = 0; j < g.N - 2; j++)
        for (int i = 0; i < g.M; i++)
            for (int k = 0; k < 4; k++) c.slot(i, j, S_UU)[k] = U[g.c(i, j + 1) * 4 + k];
}
// MUSCL-Hancock predictor, MHM_Scheme.cpp:61-96 (x) / :148-183 (y): both face states
// receive 0.5*(F(U_lo)-F(U_hi))*dt/h, where F uses the states before the update.
void predictor(Ctx& c, double dt, int ydir) {
    const Grid& g = c.g;
    double h = ydir ? c.dy : c.dx;
    int s_lo = ydir ? S_UD : S_UL, s_hi = ydir ? S_UU : S_UR;
    int f_lo = ydir ? S_GD : S_FL, f_hi = ydir ? S_GU : S_FR;
    if (ydir) { conserve2flux_y(c, s_lo, f_lo); conserve2flux_y(c, s_hi, f_hi); }
    else { conserve2flux_x(c, s_lo, f_lo); conserve2flux_x(c, s_hi, f_hi); }
    if (!ydir) {
#pragma omp parallel for
        for (int j = 0; j < g.N; j++) {
            for (int i = g.ng - 1; i < g.M - 1; i++)
                for (int k = 0; k < 4; k++)
                    c.slot(i, j, s_lo)[k] = c.slot(i, j, s_lo)[k] + 0.5 * (c.slot(i, j, f_lo)[k] - c.slot(i, j, f_hi)[k]) * dt / h;
            for (int i = 0; i < g.M - 2; i++)
                for (int k = 0; k < 4; k++)
                    c.slot(i, j, s_hi)[k] = c.slot(i, j, s_hi)[k] + 0.5 * (c.slot(i, j, f_lo)[k] - c.slot(i, j, f_hi)[k]) * dt / h;
        }
    } else {
#pragma omp parallel for
        for (int j = g.ng - 1; j < g.N - 1; j++)
            for (int i = 0; i < g.M; i++)
                for (int k = 0; k < 4; k++)
                    c.slot(i, j, s_lo)[k] = c.slot(i, j, s_lo)[k] + 0.5 * (c.slot(i, j, f_lo)[k] - c.slot(i, j, f_hi)[k]) * dt / h;
#pragma omp parallel for
        for (int j = 0; j < g.N - 2; j++)
            for (int i = 0; i < g.M; i++)
                for (int k = 0; k < 4; k++)
                    c.slot(i, j, s_hi)[k] = c.slot(i, j, s_hi)[k] + 0.5 * (c.slot(i, j, f_lo)[k] - c.slot(i, j, f_hi)[k]) * dt / h;
    }
}
// conservative updates, MUSCL_Scheme.cpp:118-132 (x) and :180-194 (y)
void update_x(Ctx& c, double dt, const double* U, double* Un, const double* XY) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++) {
            const double t = XY[g.c(i, j) * 6 + 5];
            if (t == 0) {
                for (int k = 0; k < 4; k++)
                    Un[g.c(i, j) * 4 + k] = U[g.c(i, j) * 4 + k] -
                        dt * (c.slot(i, j, S_FOLD)[k] - c.slot(i - 1, j, S_FOLD)[k]) / c.dx;
            } else if (t == 3 && g_cut) {  // cut cell: x part of the aperture-weighted balance
                const double* cc = g_cut + g.c(i, j) * 8;
                for (int k = 0; k < 4; k++)
                    Un[g.c(i, j) * 4 + k] = U[g.c(i, j) * 4 + k] -
                        dt * ((cc[2] * c.slot(i, j, S_FOLD)[k] - cc[1] * c.slot(i - 1, j, S_FOLD)[k]) * c.dy) /
                            (cc[0] * c.dx * c.dy);
            }
        }
}
void update_y(Ctx& c, double dt, const double* U, double* Un, const double* XY) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++) {
            const double t = XY[g.c(i, j) * 6 + 5];
            if (t == 0) {
                for (int k = 0; k < 4; k++)
                    Un[g.c(i, j) * 4 + k] = Un[g.c(i, j) * 4 + k] -
                        dt * (c.slot(i, j, S_GOLD)[k] - c.slot(i, j - 1, S_GOLD)[k]) / c.dy;
            } else if (t == 3 && g_cut) {  // y part and the wall's pressure force
                const double* cc = g_cut + g.c(i, j) * 8;
                const double* q = U + g.c(i, j) * 4;
                const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0];
                const double p = (c.gam - 1) * (q[3] - 0.5 * rho * (u * u + v * v));
                const double wall[4] = {0.0, p * cc[5], p * cc[6], 0.0};
                for (int k = 0; k < 4; k++)
                    Un[g.c(i, j) * 4 + k] = Un[g.c(i, j) * 4 + k] -
                        dt * ((cc[4] * c.slot(i, j, S_GOLD)[k] - cc[3] * c.slot(i, j - 1, S_GOLD)[k]) * c.dx +
                              wall[k]) / (cc[0] * c.dx * c.dy);
            }
        }
}
// Lax_Scheme.cpp:27-62 (16-double record: U_TMP slot 0, F_OLD slot 2, G_OLD slot 3)
void advance_lax(Ctx& c, double dt, const double* U, double* Un, const double* XY) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++) {
            double* t = c.slot(i, j, 0);
            for (int k = 0; k < 4; k++) t[k] = U[g.c(i, j) * 4 + k];
            flux_x(c.gam, t, c.slot(i, j, 2));
            flux_y(c.gam, t, c.slot(i, j, 3));
        }
#pragma omp parallel for
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 0)
                for (int k = 0; k < 4; k++)
                    Un[g.c(i, j) * 4 + k] =
                        0.25 * (U[g.c(i - 1, j) * 4 + k] + U[g.c(i + 1, j) * 4 + k] + U[g.c(i, j + 1) * 4 + k] + U[g.c(i, j - 1) * 4 + k]) -
                        0.5 * dt * (c.slot(i + 1, j, 2)[k] - c.slot(i - 1, j, 2)[k]) / c.dx -
                        0.5 * dt * (c.slot(i, j + 1, 3)[k] - c.slot(i, j - 1, 3)[k]) / c.dy;
}

// Corrected mode only (no reference counterpart): a face whose reconstructed state
// has rho <= 0 or p <= 0 on either side falls back to the two cell averages.
inline bool physical(double gam, const double* q) {
    double rho = q[0], u = q[1] / rho, v = q[2] / rho;
    double p = (gam - 1) * (q[3] - 0.5 * rho * (u * u + v * v));
    return rho > 0 && p > 0;
}
void positivity_fallback(Ctx& c, const double* U, int ydir) {
    const Grid& g = c.g;
    int s_lo = ydir ? S_UD : S_UL, s_hi = ydir ? S_UU : S_UR;
#pragma omp parallel for
    for (int j = 0; j < g.N - (ydir ? 1 : 0); j++)
        for (int i = 0; i < g.M - (ydir ? 0 : 1); i++) {
            double* lo = c.slot(i, j, s_lo);
            double* hi = c.slot(i, j, s_hi);
            if (physical(c.gam, lo) && physical(c.gam, hi)) continue;
            const double* a = U + g.c(i, j) * 4;
            const double* b = U + (ydir ? g.c(i, j + 1) : g.c(i + 1, j)) * 4;
            for (int k = 0; k < 4; k++) { lo[k] = a[k]; hi[k] = b[k]; }
        }
}

// Corrected mode only (no reference counterpart): density and pressure floors on
// updated fluid cells, so near-vacuum regions behind bodies stay computable.
const double RHO_FLOOR = 1e-8, P_FLOOR = 1e-8;
inline void apply_floor(double gam, double* q) {
    if (!(q[0] > RHO_FLOOR)) q[0] = RHO_FLOOR;
    double u = q[1] / q[0], v = q[2] / q[0];
    double ke = 0.5 * q[0] * (u * u + v * v);
    double p = (gam - 1) * (q[3] - ke);
    if (!(p > P_FLOOR)) q[3] = P_FLOOR / (gam - 1) + ke;
}
void floors(Ctx& c, double* Un, const double* XY) {
    const Grid& g = c.g;
#pragma omp parallel for
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 0 ||
                (XY[g.c(i, j) * 6 + 5] == 3 && !(g_cut && g_cut[g.c(i, j) * 8] < 0.5)))  // small cells: after merging
                apply_floor(c.gam, Un + g.c(i, j) * 4);
}

void advance(Ctx& c, int scheme, int rs, int lim, double dt, const double* U, double* Un, const double* XY) {
    if (scheme == SCH_LAX) { advance_lax(c, dt, U, Un, XY); return; }
    if (scheme == SCH_GODUNOV) copy_x(c, U);
    else recon_x(c, U, lim, scheme == SCH_MHM ? c.g.ng - 1 : 1);
    if (scheme == SCH_MHM) predictor(c, dt, 0);
    if (!c.quirks && scheme != SCH_GODUNOV) positivity_fallback(c, U, 0);
    riemann_x(c, rs);
    update_x(c, dt, U, Un, XY);
    if (scheme == SCH_GODUNOV) copy_y(c, U);
    else recon_y(c, U, lim);
    if (scheme == SCH_MHM) predictor(c, dt, 1);
    if (!c.quirks && scheme != SCH_GODUNOV) positivity_fallback(c, U, 1);
    riemann_y(c, rs);
    update_y(c, dt, U, Un, XY);
    if (!c.quirks) floors(c, Un, XY);
}

inline void cons_to_prim(double gam, const double* U, double& rho, double& u, double& v, double& p) {
    rho = U[0]; u = U[1] / rho; v = U[2] / rho;
    p = (gam - 1) * (U[3] - 0.5 * rho * (u * u + v * v));
}
inline void prim_to_cons(double gam, double rho, double u, double v, double p, double* U) {
    U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = p / (gam - 1) + 0.5 * rho * (u * u + v * v);
}

}  // namespace

extern "C" {

void or_set_root_mode(int mode) { g_root_mode = mode; }
void or_set_pvrs_stray(int on) { g_pvrs_stray = on; }
// counts of the PVRS y solver's stray-store branch since the last call (reset on read)
void or_pvrs_stray_counts(long long* faces, long long* gold) {
    *faces = g_stray_faces; *gold = g_stray_gold;
    g_stray_faces = g_stray_gold = 0;
}
int or_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// scratch record width for a scheme: EQDefine.H (7 slots of 4, Lax 4 slots of 4)
int or_tmp_width(int scheme) { return scheme == SCH_LAX ? 16 : 28; }

double or_limiter(int lim, int quirks, double a, double P, double Q) { return limit(lim, quirks, a, P, Q); }

void or_flux(int ydir, double gam, const double* U, double* F) {
    if (ydir) flux_y(gam, U, F); else flux_x(gam, U, F);
}

// one face flux from two conserved states (lo = left/lower), for unit tests
void or_face_flux(int rs, int ydir, double gam, const double* Ulo, const double* Uhi, double* out) {
    Prim a = ydir ? prim_y(gam, Ulo) : prim_x(gam, Ulo);
    Prim b = ydir ? prim_y(gam, Uhi) : prim_x(gam, Uhi);
    if (rs == RS_SW) {
        double f[4]; sw_face(gam, a, b, ydir, f);
        if (ydir) { out[0] = f[0]; out[1] = f[2]; out[2] = f[1]; out[3] = f[3]; }
        else for (int k = 0; k < 4; k++) out[k] = f[k];
    } else if (rs == RS_PVRS) {
        PvrsOut o; pvrs_face(gam, Ulo, Uhi, a, b, ydir, o);
        if (ydir) { out[0] = o.f[0]; out[1] = o.f[2]; out[2] = o.f[1]; out[3] = o.f[3]; }
        else for (int k = 0; k < 4; k++) out[k] = o.f[k];
    } else {
        double Flo[4], Fhi[4], SL, SR;
        if (ydir) { flux_y(gam, Ulo, Flo); flux_y(gam, Uhi, Fhi); }
        else { flux_x(gam, Ulo, Flo); flux_x(gam, Uhi, Fhi); }
        hll_speeds(gam, a, b, SL, SR);
        for (int k = 0; k < 4; k++) {
            if (SL >= 0) out[k] = Flo[k];
            else if (SR <= 0) out[k] = Fhi[k];
            else out[k] = (SR * Flo[k] - SL * Fhi[k] + SR * SL * (Uhi[k] - Ulo[k])) / (SR - SL);
        }
    }
}

// The HLL wave-speed estimates of one face (S[0] = S_lo, S[1] = S_hi), in the
// reference form (quirks = 1) or the corrected one (quirks = 0), for unit tests.
void or_hll_speeds(int quirks, int ydir, double gam, const double* Ulo, const double* Uhi, double* S) {
    Prim a = ydir ? prim_y(gam, Ulo) : prim_x(gam, Ulo);
    Prim b = ydir ? prim_y(gam, Uhi) : prim_x(gam, Uhi);
    hll_speeds(gam, a, b, S[0], S[1], quirks);
}

// Riemann_solverX / Riemann_solverY (Riemann_solver.H:22-42) on a whole scratch
// record T (M*N cells x 28 doubles), as MUSCL/MHM/Godunov call them; the PVRS
// stray store follows or_set_pvrs_stray.
void or_riemann_record(int rs, int ydir, int quirks, int nx, int ny, int ng, double gam, double* T) {
    Ctx c(nx, ny, ng, 1.0, 1.0, gam, 28, T, quirks);
    if (ydir) riemann_y(c, rs); else riemann_x(c, rs);
}

// Conserve2Flux / Conserve2FluxX / Conserve2FluxY (Conserve2Flux.cpp:4-112): n
// states at a stride of `stride` doubles, x flux to F, y flux to G (either null).
void or_flux_record(double gam, const double* U, double* F, double* G, long long n, int stride) {
    for (long long c = 0; c < n; c++) {
        if (F) flux_x(gam, U + c * stride, F + c * stride);
        if (G) flux_y(gam, U + c * stride, G + c * stride);
    }
}

// Advance (one scheme pass) -- *_Scheme.cpp. T must hold or_tmp_width(scheme)
// doubles per cell and persists between calls like the reference's U_TMP.
void or_advance(int scheme, int rs, int lim, int quirks, double L, double H, int nx, int ny, int ng,
                double gam, double dt, const double* U_OLD, double* T, double* U_NEW, const double* XY) {
    Ctx c(nx, ny, ng, L, H, gam, or_tmp_width(scheme), T, quirks);
    advance(c, scheme, rs, lim, dt, U_OLD, U_NEW, XY);
}

// ComputeDt.cpp:23-49
double or_compute_dt(double L, double H, int nx, int ny, int ng, double gam, double cfl,
                     const double* U, const double* XY) {
    Grid g(nx, ny, ng);
    double dx = L / nx;
    (void)H;
    double a_max = 0.0, u_max = 0.0, v_max = 0.0;
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++)
            if (XY[g.c(i, j) * 6 + 5] < 1 || XY[g.c(i, j) * 6 + 5] == 3) {
                const double* q = U + g.c(i, j) * 4;
                double rho = q[0], u = q[1] / q[0], v = q[2] / q[0];
                double p = (gam - 1) * (q[3] - 0.5 * rho * (u * u + v * v));
                double a = root(gam * p / rho);
                a_max = std::max(a_max, a);
                u_max = std::max(u_max, std::fabs(u));
                v_max = std::max(v_max, std::fabs(v));
            }
    return cfl * dx / (a_max + std::max(u_max, v_max));
}

// CaculateMass.cpp:22-38 (sequential order)
double or_mass(double L, double H, int nx, int ny, int ng, const double* U, const double* XY) {
    Grid g(nx, ny, ng);
    double mass = 0, dx = L / nx, dy = H / ny;
    for (int j = 0; j < g.N; j++)
        for (int i = ng; i < nx + ng; i++)
            if (XY[g.c(i, j) * 6 + 5] < 0.5) mass += U[g.c(i, j) * 4] * dx * dy;
            else if (XY[g.c(i, j) * 6 + 5] == 3 && g_cut) mass += g_cut[g.c(i, j) * 8] * U[g.c(i, j) * 4] * dx * dy;
    return mass;
}

// Psy_coord.cpp:19-37 (square-cell spacing for y, as in the reference)
void or_psy_coord(double lo_x, double lo_y, double L, double H, int nx, int ny, int ng, double* XY) {
    Grid g(nx, ny, ng);
    (void)H;
    const double dx = L / nx, dy = L / nx;
    double Li = -dx * ng + lo_x + dx / 2;
    double Hi = -dy * ng + lo_y + dy / 2;
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++) {
            XY[g.c(i, j) * 6 + 0] = Li + i * dx;
            XY[g.c(i, j) * 6 + 1] = Hi + j * dx;
        }
}

// Level-set hooks. kind 0: phi = 1 (Test/Sod/Level_Set_function.cpp:26-35);
// 1: inclined channel, prm = {angle_deg, AxisD} (Test/InclineSod/Level_Set_function.cpp:8-56);
// 2: annulus, prm = {f0, f1} (Test/Vortex/Level_Set_function.cpp:15-40);
// 3: phi given per cell in `phi`.
void or_level_set_hook(int kind, const double* prm, int nx, int ny, int ng, double* XY, const double* phi) {
    Grid g(nx, ny, ng);
    if (kind == 1) {
        double angle = 3.1415926535 * prm[0] / 180;
        double Axisd = prm[1];
        double AxisD = Axisd / std::cos(angle);
        double k = std::tan(angle);
        double nrm = std::pow(1 + k * k, 0.5);
        auto line = [&](double x, double y, double b, int dir) {
            if (y > k * x + b) return dir * std::fabs(k * x - y + b) / nrm;
            return -dir * std::fabs(k * x - y + b) / nrm;
        };
        for (int j = 0; j < g.N; j++)
            for (int i = 0; i < g.M; i++) {
                double x = XY[g.c(i, j) * 6], y = XY[g.c(i, j) * 6 + 1];
                double p1 = line(x, y, 0.05, 1), p2 = line(x, y, AxisD, -1);
                XY[g.c(i, j) * 6 + 2] = (p1 < p2) ? p1 : p2;
            }
    } else if (kind == 2) {
        const double rv = 0.40;
        double rl0 = prm[0] * rv, rl1 = prm[1] * rv;
        for (int j = 0; j < g.N; j++)
            for (int i = 0; i < g.M; i++) {
                double xp = XY[g.c(i, j) * 6], yp = XY[g.c(i, j) * 6 + 1];
                double r = std::sqrt((xp - 0.50) * (xp - 0.50) + (yp - 0.50) * (yp - 0.50));
                double p1 = rl1 - r, p2 = r - rl0;
                XY[g.c(i, j) * 6 + 2] = (p1 < p2) ? p1 : p2;
            }
    } else if (kind == 3) {
        for (size_t q = 0; q < (size_t)g.M * g.N; q++) XY[q * 6 + 2] = phi[q];
    } else {
        for (size_t q = 0; q < (size_t)g.M * g.N; q++) XY[q * 6 + 2] = 1;
    }
}

// Level_Set.cpp:71-158: normals by central differences, fluid/solid split,
// solid cells within ng of a fluid cell along an axis become ghosts, domain ring = ghost.
void or_classify(int nx, int ny, int ng, double* XY) {
    Grid g(nx, ny, ng);
    auto at = [&](int i, int j, int s) -> double& { return XY[g.c(i, j) * 6 + s]; };
    for (int j = 1; j < g.N - 1; j++)
        for (int i = 1; i < g.M - 1; i++) {
            at(i, j, 3) = (at(i + 1, j, 2) - at(i - 1, j, 2)) / ((at(i + 1, j, 0) - at(i - 1, j, 0)));
            at(i, j, 4) = (at(i, j + 1, 2) - at(i, j - 1, 2)) / ((at(i, j + 1, 1) - at(i, j - 1, 1)));
        }
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++) at(i, j, 5) = (at(i, j, 2) >= 0) ? 0 : 1;
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++)
            if (at(i, j, 5) == 0)
                for (int k = 1; k < ng + 1; k++) {
                    at(i + k, j, 5) = 2 * at(i + k, j, 5);
                    at(i - k, j, 5) = 2 * at(i - k, j, 5);
                    at(i, j + k, 5) = 2 * at(i, j + k, 5);
                    at(i, j - k, 5) = 2 * at(i, j - k, 5);
                }
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++)
            if (at(i, j, 5) > 1) at(i, j, 5) = 2;
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++)
            if (i < ng || i >= nx + ng || j < ng || j >= ny + ng) at(i, j, 5) = 2;
}

// Corrected mode only (no reference counterpart): unit normals. The reference's
// normals are the unnormalised central differences of phi, which the mirror
// points and the velocity reflection use as if they were unit vectors; for a
// level set that is not a signed distance (e.g. the mask's distance to edge
// pixels) |n|^2 reaches 2 and the reflection then amplifies the ghost velocity
// every step.
void or_unit_normals(int nx, int ny, int ng, double* XY) {
    Grid g(nx, ny, ng);
    for (size_t q = 0; q < (size_t)g.M * g.N; q++) {
        double a = XY[q * 6 + 3], b = XY[q * 6 + 4];
        double r = std::sqrt(a * a + b * b);
        if (r > 0) { XY[q * 6 + 3] = a / r; XY[q * 6 + 4] = b / r; }
    }
}

// ------------------------------------------------------------------ cut cells
// No reference counterpart (the reference has no cut-cell boundary). Geometry
// from the cell-centred level set: corner values are the means of the four
// surrounding cells' phi; a face's aperture is the fluid (phi >= 0) fraction of
// the segment between its corners under linear interpolation; the volume
// fraction alpha is the area of the fluid polygon cut from the cell by those
// crossings (shoelace); the wall vector is S = -((aE - aW) dy, (aN - aS) dx),
// which closes the cell's faces so that a uniform pressure exerts no net force.
// Types: 0 all corners fluid (alpha = 1), 3 cut (0 < alpha < 1), 1 no fluid,
// 2 an inactive cell within ng of an active one (filled like an IDW ghost) and
// the domain ring. A cut cell with alpha < 1/2 is merged with a neighbour with
// alpha >= 1/2: along the normal's dominant axis first, then the other axis,
// then the opposite directions; with none it becomes inactive. Finally every
// face between an active and an inactive interior cell is closed.
static double cut_frac(double p, double q) {
    if (p >= 0 && q >= 0) return 1.0;
    if (p < 0 && q < 0) return 0.0;
    return p >= 0 ? p / (p - q) : q / (q - p);
}
void or_set_cut(const double* cut) { g_cut = cut; }
void or_cut_geometry(int nx, int ny, int ng, double L, double H, double* XY, double* cut) {
    Grid g(nx, ny, ng);
    const double dx = L / nx, dy = H / ny;
    auto phi = [&](int i, int j) { return XY[g.c(i, j) * 6 + 2]; };
    auto corner = [&](int i, int j) {  // corner (i+1/2, j+1/2)
        return 0.25 * (((phi(i, j) + phi(i + 1, j)) + phi(i, j + 1)) + phi(i + 1, j + 1));
    };
    auto interior = [&](int i, int j) { return i >= ng && i < nx + ng && j >= ng && j < ny + ng; };
    for (size_t q = 0; q < (size_t)g.M * g.N * 8; q++) cut[q] = 0.0;
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++) {
            double* cc = cut + g.c(i, j) * 8;
            cc[7] = -1;
            if (!interior(i, j)) { XY[g.c(i, j) * 6 + 5] = 2; continue; }
            const double sw = corner(i - 1, j - 1), se = corner(i, j - 1), ne = corner(i, j), nw = corner(i - 1, j);
            const double aw = cut_frac(sw, nw), ae = cut_frac(se, ne), as = cut_frac(sw, se), an = cut_frac(nw, ne);
            double alpha;
            if (sw >= 0 && se >= 0 && ne >= 0 && nw >= 0) alpha = 1.0;
            else if (sw < 0 && se < 0 && ne < 0 && nw < 0) alpha = 0.0;
            else {
                // polygon around the unit square SW(0,0) SE(1,0) NE(1,1) NW(0,1)
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
            cc[5] = -((ae - aw) * dy);
            cc[6] = -((an - as) * dx);
            XY[g.c(i, j) * 6 + 5] = alpha >= 1.0 ? 0 : (alpha > 0.0 ? 3 : 1);
        }
    // small cells choose a merge target
    for (int j = ng; j < ny + ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            double* cc = cut + g.c(i, j) * 8;
            if (XY[g.c(i, j) * 6 + 5] != 3 || cc[0] >= 0.5) continue;
            const double n_x = XY[g.c(i, j) * 6 + 3], n_y = XY[g.c(i, j) * 6 + 4];
            const int sx = n_x >= 0 ? 1 : -1, sy = n_y >= 0 ? 1 : -1;
            int ci[4], cj[4];
            if (std::fabs(n_x) >= std::fabs(n_y)) {
                ci[0] = i + sx; cj[0] = j; ci[1] = i; cj[1] = j + sy; ci[2] = i - sx; cj[2] = j; ci[3] = i; cj[3] = j - sy;
            } else {
                ci[0] = i; cj[0] = j + sy; ci[1] = i + sx; cj[1] = j; ci[2] = i; cj[2] = j - sy; ci[3] = i - sx; cj[3] = j;
            }
            for (int q = 0; q < 4; q++) {
                if (!interior(ci[q], cj[q])) continue;
                const double tq = XY[g.c(ci[q], cj[q]) * 6 + 5];
                if ((tq == 0 || tq == 3) && cut[g.c(ci[q], cj[q]) * 8] >= 0.5) {
                    cc[7] = (double)g.c(ci[q], cj[q]);
                    break;
                }
            }
            if (cc[7] < 0) XY[g.c(i, j) * 6 + 5] = 1;  // no partner: inactive
        }
    // inactive cells within ng of an active cell along an axis are ghosts
    std::vector<double> t5(g.M * g.N);
    for (size_t q = 0; q < t5.size(); q++) t5[q] = XY[q * 6 + 5];
    for (int j = ng; j < ny + ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            if (t5[g.c(i, j)] != 1) continue;
            bool near = false;
            for (int k = 1; k <= ng && !near; k++) {
                const int ci[4] = {i - k, i + k, i, i}, cj[4] = {j, j, j - k, j + k};
                for (int q = 0; q < 4; q++)
                    if (interior(ci[q], cj[q]) && (t5[g.c(ci[q], cj[q])] == 0 || t5[g.c(ci[q], cj[q])] == 3))
                        near = true;
            }
            if (near) XY[g.c(i, j) * 6 + 5] = 2;
        }
    // a face between an active cell and an inactive interior one (zero-area
    // polygons where corners are exactly 0, small cells without a partner) is a
    // wall: aperture 0 and the wall vector closed again; a full cell with such a
    // face takes the cut-cell update (type 3, alpha = 1), so no flux leaves the
    // fluid through it
    for (size_t q = 0; q < t5.size(); q++) t5[q] = XY[q * 6 + 5];
    for (int j = ng; j < ny + ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            const double ti = t5[g.c(i, j)];
            if (ti != 0 && ti != 3) continue;
            double* cc = cut + g.c(i, j) * 8;
            const int ni[4] = {i - 1, i + 1, i, i}, nj[4] = {j, j, j - 1, j + 1};
            bool closed = false;
            for (int q = 0; q < 4; q++) {
                if (!interior(ni[q], nj[q])) continue;
                const double tq = t5[g.c(ni[q], nj[q])];
                if (tq == 0 || tq == 3 || cc[1 + q] == 0.0) continue;
                cc[1 + q] = 0.0;
                closed = true;
            }
            if (!closed) continue;
            cc[5] = -((cc[2] - cc[1]) * dy);
            cc[6] = -((cc[4] - cc[3]) * dx);
            if (ti == 0) {
                cc[0] = 1.0;
                XY[g.c(i, j) * 6 + 5] = 3;
            }
        }
}

// Small-cell merging after an update: each target and the small cells that
// chose it (neighbours W, E, S, N in that order) take the alpha-weighted mean.
void or_cut_merge(int nx, int ny, int ng, double gam, int quirks, const double* XY, const double* cut, double* U) {
    Grid g(nx, ny, ng);
    std::vector<double> out(U, U + (size_t)g.M * g.N * 4);
    for (int j = ng; j < ny + ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            const size_t t = g.c(i, j);
            const double tt = XY[t * 6 + 5];
            if (!((tt == 0 || tt == 3) && cut[t * 8] >= 0.5)) continue;
            const size_t nb[4] = {g.c(i - 1, j), g.c(i + 1, j), g.c(i, j - 1), g.c(i, j + 1)};
            size_t mem[4];
            int n = 0;
            for (int q = 0; q < 4; q++)
                if (XY[nb[q] * 6 + 5] == 3 && cut[nb[q] * 8 + 7] == (double)t) mem[n++] = nb[q];
            if (!n) continue;
            double asum = cut[t * 8];
            for (int q = 0; q < n; q++) asum = asum + cut[mem[q] * 8];
            double v[4];
            for (int k = 0; k < 4; k++) {
                v[k] = cut[t * 8] * U[t * 4 + k];
                for (int q = 0; q < n; q++) v[k] = v[k] + cut[mem[q] * 8] * U[mem[q] * 4 + k];
                v[k] = v[k] / asum;
            }
            if (!quirks) apply_floor(gam, v);  // the small cells' provisional states were not floored
            for (int k = 0; k < 4; k++) {
                out[t * 4 + k] = v[k];
                for (int q = 0; q < n; q++) out[mem[q] * 4 + k] = v[k];
            }
        }
    std::copy(out.begin(), out.end(), U);
}

}  // extern "C"

namespace {
// Mirror_IDX.cpp:20-62: lower-left donor index of the point stored at slot k of a
// W-wide record, by half-open interval search along the row / column 0.
static void mirror_idx(const Grid& g, double* S, int W, const double* XY, int k) {
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 2) {
                double xm = S[g.c(i, j) * W + k];
                for (int ii = 0; ii < g.M - 1; ii++)
                    if (xm < XY[g.c(ii + 1, j) * 6] && xm >= XY[g.c(ii, j) * 6]) S[g.c(i, j) * W + k + 2] = ii;
            }
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 2) {
                double ym = S[g.c(i, j) * W + k + 1];
                for (int jj = 0; jj < g.N - 1; jj++)
                    if (ym < XY[g.c(0, jj + 1) * 6 + 1] && ym >= XY[g.c(0, jj) * 6 + 1]) S[g.c(i, j) * W + k + 3] = jj;
            }
}
// inverse-distance weights over the 4 donors, IDWBaseGFM/GFM_Method.cpp:52-110;
// `keep(type)` picks the donors that get a weight.
template <class Keep>
static void idw_weights(const Grid& g, double* S, int W, const double* XY, int kp, int ka, Keep keep) {
    for (int j = g.ng; j < g.N - g.ng; j++)
        for (int i = g.ng; i < g.nx + g.ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 2) {
                double* s = S + g.c(i, j) * W;
                int IDX = (int)s[kp + 2], IDY = (int)s[kp + 3];
                const int di[4] = {0, 0, 1, 1}, dj[4] = {0, 1, 1, 0};
                for (int q = 0; q < 4; q++) {
                    const double* d = XY + g.c(IDX + di[q], IDY + dj[q]) * 6;
                    double a = 0;
                    if (keep(d[5])) {
                        double dd = std::sqrt((d[0] - s[kp]) * (d[0] - s[kp]) + (d[1] - s[kp + 1]) * (d[1] - s[kp + 1]));
                        a = 1 / dd;
                    }
                    s[ka + q] = a;
                }
            }
}

}  // namespace

extern "C" {

int or_sch_width(int bc) {
    switch (bc) {
        case BC_IDW: case BC_CUT: return 8;
        case BC_LINEAR: case BC_RIEMANN: return 4;
        case BC_BCGFM: case BC_HBCGFM: return 16;
        case BC_SIMPLE: return 2;
        default: return 1;
    }
}

// Scheme_Index: IDWBaseGFM/GFM_Method.cpp:5-113, LinearBaseGFM/GFM_Method.cpp:5-47,
// BCGFM/BCGFM_Method.cpp:7-210, HBCGFM/HBCGFM_Method.cpp:56-91.
void or_scheme_index(int bc, int nx, int ny, int ng, const double* XY, double* S) {
    Grid g(nx, ny, ng);
    int W = or_sch_width(bc);
    if (bc == BC_REGULAR) return;
    if (bc == BC_SIMPLE) {
        // SimpleBound/Scheme_Index.cpp:12-66: per ghost (type > 1), the nearest
        // fluid cell (type < 1) within ng cells along -x (else +x) and along -y
        // (else +y); k steps away gives the mirrored donor offset 1 - 2k (2k - 1)
        auto t = [&](int i, int j) { return XY[g.c(i, j) * 6 + 5]; };
        for (int j = ng; j < g.N - ng; j++)
            for (int i = ng; i < nx + ng; i++) {
                double* s = S + g.c(i, j) * W;
                s[0] = 0; s[1] = 0;
                if (!(t(i, j) > 1)) continue;
                for (int k = 0; k < ng + 1; k++)
                    if (t(i - k, j) < 1) { s[0] = 1 - 2 * k; break; }
                if (s[0] == 0)
                    for (int k = 0; k < ng + 1; k++)
                        if (t(i + k, j) < 1) { s[0] = 2 * k - 1; break; }
                for (int k = 0; k < ng + 1; k++)
                    if (t(i, j - k) < 1) { s[1] = 1 - 2 * k; break; }
                if (s[1] == 0)
                    for (int k = 0; k < ng + 1; k++)
                        if (t(i, j + k) < 1) { s[1] = 2 * k - 1; break; }
            }
        return;
    }
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            const double* p = XY + g.c(i, j) * 6;
            double* s = S + g.c(i, j) * W;
            bool ghost = p[5] == 2;
            if (bc == BC_IDW || bc == BC_CUT || bc == BC_LINEAR || bc == BC_RIEMANN) {
                if (ghost) { s[0] = p[0] + 2 * p[3] * std::fabs(p[2]); s[1] = p[1] + 2 * p[4] * std::fabs(p[2]); }
                else { s[0] = 0; s[1] = 0; }
            } else {
                if (ghost) {
                    double dx = XY[g.c(1, 1) * 6] - XY[g.c(0, 1) * 6];
                    double dy = XY[g.c(1, 2) * 6 + 1] - XY[g.c(1, 1) * 6 + 1];
                    double delta = std::sqrt(dx * dx + dy * dy) / 2;
                    double ph = std::fabs(p[2]);
                    if (bc == BC_HBCGFM && ph > delta) {
                        s[0] = p[0] + p[3] * (2 * ph); s[1] = p[1] + p[4] * (2 * ph);
                        s[8] = p[0] + p[3] * (2 * ph); s[9] = p[1] + p[4] * (2 * ph);
                    } else {
                        s[0] = p[0] + p[3] * (ph + delta); s[1] = p[1] + p[4] * (ph + delta);
                        s[8] = p[0] + p[3] * (ph + 2.0 * delta); s[9] = p[1] + p[4] * (ph + 2.0 * delta);
                    }
                } else { s[0] = 0; s[1] = 0; s[8] = 0; s[9] = 0; }
            }
        }
    mirror_idx(g, S, W, XY, 0);
    if (bc == BC_IDW) idw_weights(g, S, W, XY, 0, 4, [](double t) { return t < 1; });
    if (bc == BC_CUT) idw_weights(g, S, W, XY, 0, 4, [](double t) { return t == 0 || t == 3; });
    if (bc == BC_BCGFM || bc == BC_HBCGFM) {
        mirror_idx(g, S, W, XY, 8);
        idw_weights(g, S, W, XY, 0, 4, [](double t) { return t != 1; });
        idw_weights(g, S, W, XY, 8, 12, [](double t) { return t != 1; });
    }
}

// Boundary: RegularBound.cpp:22-34; IDWBaseGFM/GFM_Bound.cpp:23-106;
// LinearBaseGFM/GFM_Bound.cpp:58-575; BCGFM/BCGFM_Bound.cpp:81-175; HBCGFM_Bound.cpp:163-178.
void or_boundary(int bc, int nx, int ny, int ng, double gam, double* U_OLD, double* U_NEW, const double* XY,
                 const double* S, int quirks) {
    Grid g(nx, ny, ng);
    int W = or_sch_width(bc);
    if (bc == BC_REGULAR) {
        for (int j = ng; j < g.N - ng; j++)
            for (int i = ng; i < nx + ng; i++)
                for (int k = 0; k < 4; k++) U_OLD[g.c(i, j) * 4 + k] = U_NEW[g.c(i, j) * 4 + k];
        return;
    }
    if (bc == BC_SIMPLE) {
        // SimpleBound/LS_SimpleBound.cpp:15-110: fluid cells U_OLD <- U_NEW, then each
        // ghost from its x and/or y donor of U_OLD with the normal velocity
        // component negated (the mean of both reflections when both exist), in
        // the reference's sequential order (i outer, j inner; a ghost donor read
        // after its own fill holds the new value)
        for (int j = ng; j < g.N - ng; j++)
            for (int i = ng; i < nx + ng; i++)
                if (XY[g.c(i, j) * 6 + 5] == 0)
                    for (int k = 0; k < 4; k++) U_OLD[g.c(i, j) * 4 + k] = U_NEW[g.c(i, j) * 4 + k];
        for (int i = ng; i < nx + ng; i++)
            for (int j = ng; j < g.N - ng; j++) {
                if (!(XY[g.c(i, j) * 6 + 5] > 1)) continue;
                const double* s = S + g.c(i, j) * W;
                const int IDX = (int)s[0], IDY = (int)s[1];
                double rx = 0, ux = 0, vx = 0, px = 0, ry = 0, uy = 0, vy = 0, py = 0, rw, uw, vw, pw;
                if (IDX != 0) cons_to_prim(gam, U_OLD + g.c(i + IDX, j) * 4, rx, ux, vx, px);
                if (IDY != 0) cons_to_prim(gam, U_OLD + g.c(i, j + IDY) * 4, ry, uy, vy, py);
                if (IDX != 0 && IDY != 0) {
                    rw = 0.5 * (rx + ry); uw = 0.5 * (-ux + uy); vw = 0.5 * (vx - vy); pw = 0.5 * (px + py);
                } else if (IDX != 0) {
                    rw = rx; uw = -ux; vw = vx; pw = px;
                } else if (IDY != 0) {
                    rw = ry; uw = uy; vw = -vy; pw = py;
                } else {
                    continue;
                }
                prim_to_cons(gam, rw, uw, vw, pw, U_OLD + g.c(i, j) * 4);
            }
        return;
    }
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++)
            if (XY[g.c(i, j) * 6 + 5] == 0 || (bc == BC_CUT && XY[g.c(i, j) * 6 + 5] == 3))
                for (int k = 0; k < 4; k++) U_OLD[g.c(i, j) * 4 + k] = U_NEW[g.c(i, j) * 4 + k];
    for (size_t q = 0; q < (size_t)g.M * g.N * 4; q++) U_NEW[q] = U_OLD[q];
    double dx = XY[g.c(1, 1) * 6] - XY[g.c(0, 1) * 6];
    double dy = XY[g.c(1, 2) * 6 + 1] - XY[g.c(1, 1) * 6 + 1];
    for (int j = ng; j < g.N - ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            if (!(XY[g.c(i, j) * 6 + 5] == 2)) continue;
            const double* s = S + g.c(i, j) * W;
            double n_x = XY[g.c(i, j) * 6 + 3], n_y = XY[g.c(i, j) * 6 + 4];
            // donors 1..4 = (IDX,IDY), (IDX,IDY+1), (IDX+1,IDY+1), (IDX+1,IDY)
            auto donors = [&](int kp, double r[4], double u[4], double v[4], double p[4], double t[4]) {
                int IDX = (int)s[kp + 2], IDY = (int)s[kp + 3];
                const int di[4] = {0, 0, 1, 1}, dj[4] = {0, 1, 1, 0};
                for (int q = 0; q < 4; q++) {
                    size_t cc = g.c(IDX + di[q], IDY + dj[q]);
                    cons_to_prim(gam, U_NEW + cc * 4, r[q], u[q], v[q], p[q]);
                    t[q] = XY[cc * 6 + 5];
                }
            };
            double rw, uw, vw, pw;
            if (bc == BC_IDW || bc == BC_CUT || bc == BC_BCGFM || bc == BC_HBCGFM) {
                auto idw = [&](int kp, int ka, double& R, double& Uu, double& V, double& P) {
                    double r[4], u[4], v[4], p[4], t[4];
                    donors(kp, r, u, v, p, t);
                    double a1 = s[ka], a2 = s[ka + 1], a3 = s[ka + 2], a4 = s[ka + 3];
                    if (!quirks)  // corrected mode: a donor at the point itself (1/0 weight) is the value
                        for (int q = 0; q < 4; q++)
                            if (std::isinf(s[ka + q])) { R = r[q]; Uu = u[q]; V = v[q]; P = p[q]; return; }
                    R = (a1 * r[0] + a2 * r[1] + a3 * r[2] + a4 * r[3]) / (a1 + a2 + a3 + a4);
                    Uu = (a1 * u[0] + a2 * u[1] + a3 * u[2] + a4 * u[3]) / (a1 + a2 + a3 + a4);
                    V = (a1 * v[0] + a2 * v[1] + a3 * v[2] + a4 * v[3]) / (a1 + a2 + a3 + a4);
                    P = (a1 * p[0] + a2 * p[1] + a3 * p[2] + a4 * p[3]) / (a1 + a2 + a3 + a4);
                };
                if (bc == BC_IDW || bc == BC_CUT) {
                    if (!quirks && s[4] + s[5] + s[6] + s[7] == 0) continue;  // no fluid donor: keep the ghost
                    idw(0, 4, rw, uw, vw, pw);
                    double u_n = uw * n_x + vw * n_y;
                    double u_t = uw * n_y - vw * n_x;
                    u_n = -u_n;
                    uw = u_n * n_x + u_t * n_y;
                    vw = u_n * n_y - u_t * n_x;
                } else {
                    double delta = std::sqrt(dx * dx + dy * dy) / 2;
                    double R_E = bc == BC_HBCGFM ? 1.0 * delta : delta;
                    double R_G = std::fabs(XY[g.c(i, j) * 6 + 2]);
                    double ri, ui, vi, pi_, re, ue, ve, pe;
                    idw(0, 4, ri, ui, vi, pi_);
                    double uni = ui * n_x + vi * n_y, uti = ui * n_y - vi * n_x;
                    idw(8, 12, re, ue, ve, pe);
                    double une = ue * n_x + ve * n_y, ute = ue * n_y - ve * n_x;
                    double un, ut;
                    if (bc == BC_HBCGFM && R_G > delta) {
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
            } else {
                // Linear/RiemannBaseGFM: one donor chosen by the fluid pattern of the 4
                double r[4], u[4], v[4], p[4], t[4];
                donors(0, r, u, v, p, t);
                int IDX = (int)s[2], IDY = (int)s[3];
                double x_d = s[0] - XY[g.c(IDX, IDY) * 6];
                double y_d = s[0] - XY[g.c(IDX, IDY) * 6];  // x-mirror in both (reference)
                bool f1 = t[0] < 1, f2 = t[1] < 1, f3 = t[2] < 1, f4 = t[3] < 1;
                bool xr = x_d > dx / 2, yu = y_d > dy / 2;
                int pick = -1;  // 0..3 = donor 1..4, -1 = none (all-zero state)
                // cases 1-15 of LinearBaseGFM/GFM_Bound.cpp:123-575
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
                if (pick >= 0) { rw = r[pick]; uw = u[pick]; vw = v[pick]; pw = p[pick]; }
                else { rw = 0; uw = 0; vw = 0; pw = 0; }
                double u_n = uw * n_x + vw * n_y;
                double u_t = uw * n_y - vw * n_x;
                u_n = -u_n;
                uw = u_n * n_x + u_t * n_y;
                vw = u_n * n_y - u_t * n_x;
            }
            prim_to_cons(gam, rw, uw, vw, pw, U_OLD + g.c(i, j) * 4);
        }
}

// Edge policies. kind 0: Test/Sod/Global_Boudary.cpp:24-75 (zero-gradient x, "reflective" y
// with the reference's row choices). kind 1: periodic in x and y (no reference counterpart).
// kind 2: supersonic inflow on the left with state prm = (rho,u,v,p), zero-gradient elsewhere
// (no reference counterpart).
void or_global_boundary(int kind, const double* prm, int nx, int ny, int ng, double gam, double* U_OLD,
                        const double* U_NEW) {
    Grid g(nx, ny, ng);
    auto O = [&](int i, int j, int k) -> double& { return U_OLD[g.c(i, j) * 4 + k]; };
    auto Nw = [&](int i, int j, int k) { return U_NEW[g.c(i, j) * 4 + k]; };
    if (kind == EDGE_REFERENCE) {
        for (int j = 0; j < g.N; j++)
            for (int i = ng; i < nx + ng; i++)
                for (int k = 0; k < 4; k++) {
                    if (j < ng) O(i, j, k) = Nw(i, 2 * ng - 1, k);
                    if (j >= ny + ng) O(i, j, k) = Nw(i, ny + ng - 1, k);
                }
        for (int j = 0; j < g.N; j++)
            for (int i = ng; i < nx + ng; i++) {
                if (j < ng) O(i, j, 2) = -Nw(i, 2 * ng - j - 1, 2);
                if (j >= ny + ng) O(i, j, 2) = -Nw(i, 2 * (ny + ng) - 1 - j, 2);
            }
        for (int j = 0; j < g.N; j++)
            for (int i = 0; i < ng; i++)
                for (int k = 0; k < 4; k++) {
                    O(i, j, k) = Nw(ng, j, k);
                    O(nx + ng + i, j, k) = Nw(nx + ng - 1, j, k);
                }
    } else if (kind == EDGE_PNGGEO) {
        // Test/PNGGeo/Global_Boudary.cpp:14-79: rows as the Sod hook over the
        // interior columns, right columns zero-gradient from U_NEW, left columns
        // the fixed post-shock state prm = (rho, u, v, p); ring corners untouched
        for (int i = ng; i < nx + ng; i++)
            for (int j = 0; j < ng; j++)
                for (int k = 0; k < 4; k++) {
                    O(i, j, k) = Nw(i, 2 * ng - 1, k);
                    O(i, ny + ng + j, k) = Nw(i, ny + ng - 1, k);
                }
        for (int i = ng; i < nx + ng; i++)
            for (int j = 0; j < ng; j++) {
                O(i, j, 2) = -Nw(i, 2 * ng - j - 1, 2);
                O(i, ny + ng + j, 2) = -Nw(i, ny + ng - 1 - j, 2);
            }
        for (int j = ng; j < ny + ng; j++)
            for (int i = 0; i < ng; i++)
                for (int k = 0; k < 4; k++) {
                    O(i, j, k) = Nw(ng, j, k);
                    O(nx + ng + i, j, k) = Nw(nx + ng - 1, j, k);
                }
        double Uin[4];
        prim_to_cons(gam, prm[0], prm[1], prm[2], prm[3], Uin);
        for (int j = ng; j < ny + ng; j++)
            for (int i = 0; i < ng; i++)
                for (int k = 0; k < 4; k++) O(i, j, k) = Uin[k];
    } else if (kind == EDGE_PERIODIC) {
        for (int j = 0; j < g.N; j++)
            for (int i = 0; i < g.M; i++) {
                if (i >= ng && i < nx + ng && j >= ng && j < ny + ng) continue;
                int si = i < ng ? i + nx : (i >= nx + ng ? i - nx : i);
                int sj = j < ng ? j + ny : (j >= ny + ng ? j - ny : j);
                for (int k = 0; k < 4; k++) O(i, j, k) = Nw(si, sj, k);
            }
    } else {
        double Uin[4];
        prim_to_cons(gam, prm[0], prm[1], prm[2], prm[3], Uin);
        for (int j = 0; j < g.N; j++)
            for (int i = 0; i < g.M; i++) {
                if (i >= ng && i < nx + ng && j >= ng && j < ny + ng) continue;
                if (i < ng) { for (int k = 0; k < 4; k++) O(i, j, k) = Uin[k]; continue; }
                int si = i >= nx + ng ? nx + ng - 1 : i;
                int sj = j < ng ? ng : (j >= ny + ng ? ny + ng - 1 : j);
                for (int k = 0; k < 4; k++) O(i, j, k) = Nw(si, sj, k);
            }
    }
}

// Initialisers. kind 0: Test/Sod/Initialize.cpp:29-52 (split on the row index);
// 1: Test/InclineSod/Initialize.cpp:21-72 (prm = {angle_deg, rho_L,u_L,v_L,p_L, rho_R,u_R,v_R,p_R});
// 2: Test/Vortex/Initialize.cpp:15-98 (prm = {fa, fw});
// 3: Sod split at x = prm[0] with prm[1..8] = left/right primitive states (no reference hook);
// 4: uniform state prm = {rho,u,v,p} (no reference hook);
// 5: isentropic vortex, prm = {x0, y0, beta, rho_inf, u_inf, v_inf, p_inf} (no reference hook);
// 6: Test/PNGGeo/Initialize.cpp:14-54, a shock of Mach prm[1] at x = prm[0]: the
//    state prm[2..5] = (rho, u, v, p) where x > prm[0], its normal-shock image
//    behind (x <= prm[0]).
// prm for kind 0: {rho_L,u_L,v_L,p_L, rho_R,u_R,v_R,p_R}.
void or_initialize(int kind, const double* prm, int nx, int ny, int ng, double gam, double* U_OLD,
                   double* U_NEW, const double* XY) {
    Grid g(nx, ny, ng);
    for (int j = 0; j < g.N; j++)
        for (int i = 0; i < g.M; i++) {
            double* U = U_OLD + g.c(i, j) * 4;
            const double* p = XY + g.c(i, j) * 6;
            if (kind == 0) {
                const double* s = (j < (nx + 2 * ng) / 2) ? prm : prm + 4;
                U[0] = s[0]; U[1] = s[0] * s[1]; U[2] = s[0] * s[2];
                U[3] = s[3] / (gam - 1) + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2]);
            } else if (kind == 1) {
                double angle = 3.1415926535 * prm[0] / 180;
                if (p[5] < 0.5 || p[5] > 1.5) {
                    const double* s = (p[1] < -(1 / std::tan(angle)) * p[0] + 0.5 * (1 / std::tan(angle)) + 0.5)
                                          ? prm + 1 : prm + 5;
                    U[0] = s[0]; U[1] = s[0] * s[1]; U[2] = s[0] * s[2];
                    U[3] = s[3] / (gam - 1) + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2]);
                } else { U[0] = -1; U[1] = 0; U[2] = 0; U[3] = -1; }
            } else if (kind == 2) {
                const double rho0 = 1.0, p0 = 2.0, xlc = 0.50, ylc = 0.50, rvortex = 0.40, pi = M_PI;
                double ua = prm[0] * rvortex * pi;
                double ur = 0, pr = 0, u = 0, v = 0;
                U[0] = rho0; U[1] = rho0 * u; U[2] = rho0 * v;
                U[3] = (p0 + 2.0 * rho0 * ua * ua * pr) / (gam - 1) + 0.5 * rho0 * (u * u + v * v);
                if (p[5] < 0.5 || p[5] > 1.5) {
                    double xp = p[0], yp = p[1];
                    double r = std::pow((xp - xlc) * (xp - xlc) + (yp - ylc) * (yp - ylc), 0.5);
                    double phi = std::atan((yp - ylc) / (xp - xlc));
                    if (xp - xlc < 0.0) phi = phi + pi;
                    if (r < rvortex) {
                        if (r < rvortex / 2.0) {
                            ur = ua * 2.0 * r / rvortex;
                            pr = (r / rvortex) * (r / rvortex) + 1.0 - 2.0 * std::log(2.0);
                        } else {
                            ur = ua * 2.0 * (1.0 - r / rvortex);
                            pr = (r / rvortex) * (r / rvortex) + 3.0 - 4.0 * r / rvortex + 2.0 * std::log(r / rvortex);
                        }
                        u = -std::sin(phi) * ur;
                        v = std::cos(phi) * ur;
                    } else { pr = 0.0; u = 0.0; v = 0.0; }
                    U[0] = rho0; U[1] = rho0 * u; U[2] = rho0 * v;
                    U[3] = (p0 + 2.0 * rho0 * ua * ua * pr) / (gam - 1) + 0.5 * rho0 * (u * u + v * v);
                }
            } else if (kind == 3) {
                const double* s = (p[0] < prm[0]) ? prm + 1 : prm + 5;
                prim_to_cons(gam, s[0], s[1], s[2], s[3], U);
            } else if (kind == 4) {
                prim_to_cons(gam, prm[0], prm[1], prm[2], prm[3], U);
            } else if (kind == 6) {
                const double rho_1 = prm[2], u_1 = prm[3], v_1 = prm[4], p_1 = prm[5], Ms = prm[1];
                if (p[0] > prm[0]) {
                    U[0] = rho_1; U[1] = rho_1 * u_1; U[2] = rho_1 * v_1;
                    U[3] = p_1 / (gam - 1) + 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1);
                } else {
                    double rho_2 = rho_1 * (gam + 1) * Ms * Ms / (2 + (gam - 1) * Ms * Ms);
                    double p_2 = p_1 * ((2 * gam) * Ms * Ms / (gam + 1) - (gam - 1) / (gam + 1));
                    double u_2 = Ms * (1 - rho_1 / rho_2);
                    U[0] = rho_2; U[1] = rho_2 * u_2; U[2] = 0;
                    U[3] = p_2 / (gam - 1) + 0.5 * rho_2 * (u_2 * u_2);
                }
            } else {
                double x0 = prm[0], y0 = prm[1], beta = prm[2];
                double dxv = p[0] - x0, dyv = p[1] - y0, r2 = dxv * dxv + dyv * dyv;
                double T = 1 - (gam - 1) * beta * beta / (8 * gam * M_PI * M_PI) * std::exp(1 - r2);
                double rho = prm[3] * std::pow(T, 1 / (gam - 1));
                double pr = prm[6] / prm[3] * rho * T;  // p = rho*T in units where p_inf = rho_inf = 1
                double f = beta / (2 * M_PI) * std::exp(0.5 * (1 - r2));
                prim_to_cons(gam, rho, prm[4] - f * dyv, prm[5] + f * dxv, pr, U);
            }
            for (int k = 0; k < 4; k++) U_NEW[g.c(i, j) * 4 + k] = U[k];
        }
}

// Test/PNGGeo/Level_Set.cpp:31-149: phi per cell (reference cell order j*M+i)
// from an RGBA image (top row first), by the reference's brute-force search over
// all edge pixels. The edge-magnitude slot is set for pixels 1..W-2 x 1..H-2
// only; the reference leaves the border pixels' slot uninitialised and scans
// all pixels, which is taken here as 0 (no edge).
// With cells != null, only the ncell listed cells (reference cell order) are
// computed, into phi[0 .. ncell): the same search for a sample of a large grid.
static void mask_phi_impl(const unsigned char* rgba, int W, int H, int nx, int ny, int ng, double lo_x,
                          double lo_y, double hi_x, double hi_y, const double* XY, double* phi, const long long* cells,
                          long long ncell) {
    const double Psy_L = hi_x - lo_x, Psy_H = hi_y - lo_y;
    std::vector<double> v((size_t)W * H), g((size_t)W * H, 0.0);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const unsigned char* px = rgba + ((size_t)y * W + x) * 4;
            double R = px[0], G = px[1], B = px[2];
            v[(size_t)(H - y - 1) * W + x] = ((R + G + B) > 0) * 2 - 1;
        }
    int edg_num = 0;
    for (int y = 1; y < H - 1; y++)
        for (int x = 1; x < W - 1; x++) {
            double dx = v[(size_t)y * W + x] - v[(size_t)y * W + x - 1];
            double dy = v[(size_t)y * W + x] - v[(size_t)(y - 1) * W + x];
            g[(size_t)y * W + x] = (std::fabs(dx) > std::fabs(dy)) ? std::fabs(dx) : std::fabs(dy);
            if (g[(size_t)y * W + x] > 0.1) edg_num++;
        }
    std::vector<double> ex(edg_num), ey(edg_num);
    int k = 0;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double dxp = (Psy_L / nx) * (nx + 2 * ng) / W;
            double dyp = (Psy_H / ny) * (ny + 2 * ng) / H;
            if (g[(size_t)y * W + x] > 0.1) {
                ex[k] = -ng * (Psy_L / nx) + dxp * x;
                ey[k] = -ng * (Psy_H / ny) + dyp * y;
                k++;
            }
        }
    const int M = nx + 2 * ng, N = ny + 2 * ng;
    auto at = [&](int i, int j) {
        int x = (i + 0.5) * (W) / (M);
        int y = (j + 0.5) * (H) / (N);
        const double sgn = v[(size_t)y * W + x];
        const double X = XY[((size_t)j * M + i) * 6 + 0], Y = XY[((size_t)j * M + i) * 6 + 1];
        double min_distance = 100;
        for (int e = 0; e < edg_num; e++) {
            double d = std::sqrt((ex[e] - X) * (ex[e] - X) + (ey[e] - Y) * (ey[e] - Y));
            if (d < min_distance) min_distance = d;
        }
        int pn = sgn >= 0 ? 1 : -1;
        return pn * min_distance;
    };
    if (cells) {
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < ncell; q++) phi[q] = at((int)(cells[q] % M), (int)(cells[q] / M));
        return;
    }
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) phi[(size_t)j * M + i] = at(i, j);
}

void or_mask_phi(const unsigned char* rgba, int W, int H, int nx, int ny, int ng, double lo_x, double lo_y,
                 double hi_x, double hi_y, const double* XY, double* phi) {
    mask_phi_impl(rgba, W, H, nx, ny, ng, lo_x, lo_y, hi_x, hi_y, XY, phi, nullptr, 0);
}

void or_mask_phi_cells(const unsigned char* rgba, int W, int H, int nx, int ny, int ng, double lo_x, double lo_y,
                       double hi_x, double hi_y, const double* XY, const long long* cells, long long ncell,
                       double* phi) {
    mask_phi_impl(rgba, W, H, nx, ny, ng, lo_x, lo_y, hi_x, hi_y, XY, phi, cells, ncell);
}

// TimeAdvance. SingleStep (SingleStep/SingleStep.cpp:5-21) is one Advance.
// RK2 (RK2/RK2.cpp:12-48, which does not compile against the current headers)
// is restated as its evident intent, Heun's method: U1 = Advance(U^n);
// Boundary; Global_Boundary; U2 = Advance(U1); U_NEW = (U^n + U2) / 2 on fluid
// cells, U_NEW = U_OLD (= U1 after Boundary) elsewhere. The reference's text
// averages with U_OLD after Boundary has overwritten it with U1, which is not a
// consistent two-stage method; U^n is kept here.
void or_time_advance(int time, int scheme, int rs, int lim, int bc, int edge, const double* edge_prm, int quirks,
                     double L, double H, int nx, int ny, int ng, double gam, double dt, double* U_OLD, double* T,
                     double* U_NEW, const double* XY, const double* S) {
    if (time == 0) {
        or_advance(scheme, rs, lim, quirks, L, H, nx, ny, ng, gam, dt, U_OLD, T, U_NEW, XY);
        if (bc == BC_CUT && g_cut) or_cut_merge(nx, ny, ng, gam, quirks, XY, g_cut, U_NEW);
        return;
    }
    Grid g(nx, ny, ng);
    const size_t n4 = (size_t)g.M * g.N * 4;
    std::vector<double> U0(U_OLD, U_OLD + n4);
    or_advance(scheme, rs, lim, quirks, L, H, nx, ny, ng, gam, dt, U_OLD, T, U_NEW, XY);
    or_boundary(bc, nx, ny, ng, gam, U_OLD, U_NEW, XY, S, quirks);
    or_global_boundary(edge, edge_prm, nx, ny, ng, gam, U_OLD, U_NEW);
    std::vector<double> U2(U_NEW, U_NEW + n4);
    or_advance(scheme, rs, lim, quirks, L, H, nx, ny, ng, gam, dt, U_OLD, T, U2.data(), XY);
    for (int j = ng; j < ny + ng; j++)
        for (int i = ng; i < nx + ng; i++) {
            const size_t o = (size_t)g.c(i, j) * 4;
            const bool fluid = XY[g.c(i, j) * 6 + 5] == 0;
            for (int k = 0; k < 4; k++) U_NEW[o + k] = fluid ? 0.5 * (U0[o + k] + U2[o + k]) : U_OLD[o + k];
        }
}

// Whole-run driver in the order of Test/Sod/Sod.cpp:94-181 / Test/InclineSod/InclineSod.cpp:73-165:
// (state, geometry and scheme index already set up by the caller)
//   [if init_boundary] Boundary ; dt = ComputeDt
//   while (now_t < t_end && iter < max_iter): TimeAdvance(dt); Boundary; Global_Boundary;
//       dt = ComputeDt; mass; now_t += dt; iter++
// Returns the number of steps taken; per-step dt/time/mass go to the log arrays (may be null).
// clip = 1: a driver of the same order that shortens the last step to end at t_end
// (step_dt = min(dt, t_end - now_t)) and advances its clock by the dt it stepped
// with; TimeAdvance takes that dt (Src/Base/TimeAdvance/SingleStep/SingleStep.cpp:5-20).
int or_run(int time, int scheme, int rs, int lim, int bc, int edge, const double* edge_prm, int quirks, double L,
           double H, int nx, int ny, int ng, double gam, double cfl, double t_end, int max_iter, int init_boundary,
           double* U_OLD, double* U_NEW, const double* XY, const double* S, double* T, double* dt_log,
           double* t_log, double* mass_log, int clip) {
    if (init_boundary) or_boundary(bc, nx, ny, ng, gam, U_OLD, U_NEW, XY, S, quirks);
    double dt = or_compute_dt(L, H, nx, ny, ng, gam, cfl, U_OLD, XY);
    double now_t = 0;
    int iter = 0;
    while (now_t < t_end && iter < max_iter) {
        double step_dt = dt;
        if (clip && now_t + step_dt > t_end) step_dt = t_end - now_t;
        or_time_advance(time, scheme, rs, lim, bc, edge, edge_prm, quirks, L, H, nx, ny, ng, gam, step_dt, U_OLD,
                        T, U_NEW, XY, S);
        or_boundary(bc, nx, ny, ng, gam, U_OLD, U_NEW, XY, S, quirks);
        or_global_boundary(edge, edge_prm, nx, ny, ng, gam, U_OLD, U_NEW);
        dt = or_compute_dt(L, H, nx, ny, ng, gam, cfl, U_OLD, XY);
        double m = or_mass(L, H, nx, ny, ng, U_OLD, XY);
        now_t = now_t + (clip ? step_dt : dt);
        if (dt_log) dt_log[iter] = dt;
        if (t_log) t_log[iter] = now_t;
        if (mass_log) mass_log[iter] = m;
        iter++;
    }
    return iter;
}

}  // extern "C"
