// Per-face arithmetic of the Euler update, shared by the FP64 step kernel and
// the mixed-precision kernel (T = float for reconstruction and fluxes).
//
// Every expression keeps the reference's operation order so that, compiled
// without FP contraction, the FP64 instantiation reproduces the reference's
// results bit for bit (sound speeds use sqrt where the reference calls
// pow(x, 0.5); see the per-function citations).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define IB_HD __host__ __device__ __forceinline__

namespace ibtfo {

enum { SCH_LAX = 0, SCH_MUSCL = 1, SCH_MHM = 2, SCH_GODUNOV = 3 };
enum { RS_HLL = 0, RS_HLLC = 1, RS_PVRS = 2, RS_SW = 3 };
enum { LIM_MINMOD = 0, LIM_VANLEER = 1, LIM_VANALBADA = 2 };

template <typename T> struct FpConst;
template <> struct FpConst<double> {
    static constexpr double half_ulp_one = 0x1p-53;  // half the spacing of doubles just above 1
    static constexpr double tiny = 0x1p-968;         // below this b*half_ulp_one may round
    static constexpr double huge = 0x1.fffffffffffffp+1023;
};
template <> struct FpConst<float> {
    static constexpr float half_ulp_one = 0x1p-24f;
    static constexpr float tiny = 0x1p-100f;
    static constexpr float huge = 0x1.fffffep+127f;
};

// Square root. On the device, for x in [2^-767, 2^1023] (every sound speed and
// shock factor of a physical state), the rsq seed + Goldschmidt refinement +
// two residual corrections that the compiler's IEEE sqrt runs after its range
// scaling, without the scaling and special-value fixups: the same arithmetic on
// the same operands, so the same correctly rounded result. Other x (zero,
// tiny, infinite, negative, NaN) take the full IEEE sequence.
IB_HD double ib_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (!(x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023)) return __builtin_sqrt(x);
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
#else
    return __builtin_sqrt(x);
#endif
}
// FP32 (mixed-precision form): the hardware square root (1 ulp)
IB_HD float ib_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}
IB_HD double ib_fabs(double x) { return __builtin_fabs(x); }
IB_HD float ib_fabs(float x) { return __builtin_fabsf(x); }

// --------------------------------------------------------------- division
// Correctly rounded quotients at a fraction of the cost of the generic IEEE
// sequence (v_div_scale x2 / v_rcp / 4 fma / mul / fma / v_div_fmas / v_div_fixup):
//  * rcp_of(b): y ~ 1/b from v_rcp_f64 refined by two Newton steps (the
//    refinement the IEEE sequence itself performs);
//  * dq(a, r): q0 = a*y, residual rr = a - b*q0 (exact, one fma), q = q0 + rr*y
//    (Markstein's final correction, also the IEEE sequence's), which is RN(a/b);
//    the sign is taken from q0 so that a zero quotient keeps IEEE's sign;
//  * divisions by run constants (dx, dy, gamma, gamma-1) use y = RN(1/d)
//    computed on the host, with the same correction;
//  * one reciprocal serves all quotients with the same denominator.
// The IEEE sequence's operand scaling is omitted, so the result equals IEEE
// division for operands and quotients of magnitude in [2^-1000, 2^1000]; a zero
// or non-finite denominator gives NaN (IEEE: +-inf for a finite non-zero
// numerator), which only arises once a state is already non-physical.
template <typename T> struct Rcp { T b, y; };

IB_HD Rcp<double> rcp_of(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    return Rcp<double>{b, y};
#else
    return Rcp<double>{b, 1.0 / b};
#endif
}
// FP32 (mixed-precision form): the hardware reciprocal (1 ulp) and a plain
// product, about 2 ulp per quotient; no correct rounding is attempted there.
IB_HD Rcp<float> rcp_of(float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return Rcp<float>{b, __builtin_amdgcn_rcpf(b)};
#else
    return Rcp<float>{b, 1.0f / b};
#endif
}
IB_HD double ib_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
IB_HD float ib_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
IB_HD double ib_copysign(double a, double b) { return __builtin_copysign(a, b); }
IB_HD float ib_copysign(float a, float b) { return __builtin_copysignf(a, b); }
template <typename T> IB_HD T dq(T a, const Rcp<T>& r) {
    T q0 = a * r.y;
    T rr = ib_fma(-q0, r.b, a);
    return ib_copysign(ib_fma(rr, r.y, q0), q0);
}
IB_HD float dq(float a, const Rcp<float>& r) { return a * r.y; }
template <typename T> IB_HD T ddiv(T a, T b) { return dq(a, rcp_of(b)); }

// minmod/Slope_limiter.H:4: (a*b > 0) ? ((a/b > 1) ? b : a) : 0.
// For a*b > 0 the rounded quotient exceeds 1 exactly when |a| > |b|(1 + 2^-53)
// (the tie at 1 + 2^-53 rounds to 1.0), which is tested without a division:
// |a| - |b| is exact by Sterbenz whenever |a| < 2|b|. Tiny or infinite |b|
// falls back to the quotient itself.
template <typename T> IB_HD T lim_minmod(T a, T b) {
    if (!(a * b > T(0))) return T(0);
    T A = ib_fabs(a), B = ib_fabs(b);
    bool gt;
    if (B >= FpConst<T>::tiny && B <= FpConst<T>::huge) gt = (A >= T(2) * B) || (A - B > B * FpConst<T>::half_ulp_one);
    else gt = (a / b > T(1));
    return gt ? b : a;
}
// VanLeer/Slope_limiter.H:4 as expanded at the call sites, whose second argument
// is the unparenthesised difference P - Q: `(r*r + r) * b / (r*r + 1)` becomes
// (r*r + r)*P - Q/(r*r + 1).
template <typename T> IB_HD T lim_vanleer_ref(T a, T P, T Q) {
    T r = ddiv(a + T(1e-7), (P - Q) + T(1e-7));
    return (r > T(0)) ? ((r * r + r) * P - ddiv(Q, r * r + T(1))) : T(0);
}
// The same limiter with its argument parenthesised (corrected mode), written
// with one division: with A = a + 1e-7, B = b + 1e-7 and r = A/B,
// (r*r + r) b / (r*r + 1) = b A (A + B) / (A*A + B*B), and r > 0 iff A*B > 0.
// lim(a, b) and lim(b, a) share the denominator, so a cell's two slopes in one
// direction cost one reciprocal (vanleer_pair).
template <typename T> IB_HD T lim_vanleer_fixed(T a, T b) {
    T A = a + T(1e-7), B = b + T(1e-7);
    return (A * B > T(0)) ? ddiv(b * (A * (A + B)), A * A + B * B) : T(0);
}
// {lim(a, b), lim(b, a)} of lim_vanleer_fixed
template <typename T> IB_HD void vanleer_pair(T a, T b, T& sab, T& sba) {
    T A = a + T(1e-7), B = b + T(1e-7);
    T n = A + B;
    Rcp<T> rd = rcp_of(A * A + B * B);
    bool pos = A * B > T(0);
    sab = pos ? dq(b * (A * n), rd) : T(0);
    sba = pos ? dq(a * (B * n), rd) : T(0);
}
// VanAlbada/Slope_limiter.H:4 does not compile (undeclared r); its evident
// intent, with r defined as in the sibling limiter.
template <typename T> IB_HD T lim_vanalbada(T a, T b) {
    T r = ddiv(a + T(1e-7), b + T(1e-7));
    return (r > T(0)) ? ddiv((r + ib_fabs(r)) * b, T(1) + ib_fabs(r)) : T(0);
}
template <int LIM, typename T> IB_HD T slope(T a, T P, T Q, bool quirks) {
    if (LIM == LIM_MINMOD) return lim_minmod(a, P - Q);
    if (LIM == LIM_VANLEER) return quirks ? lim_vanleer_ref(a, P, Q) : lim_vanleer_fixed(a, P - Q);
    return lim_vanalbada(a, P - Q);
}

// Both slopes of one cell from its neighbours um, u0, up (per component):
// lo = u0 + 0.5 lim(u0 - um, up - u0) is the cell's state on its upper face
// (U_L of face +1/2), hi = u0 - 0.5 lim(up - u0, u0 - um) its state on its lower
// face (U_R of face -1/2); recon() below evaluates the same expressions face by face.
template <int LIM, typename T> IB_HD void slope_pair(const T* um, const T* u0, const T* up, T* lo, T* hi,
                                                     bool quirks) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        T sl, sh;
        if (LIM == LIM_VANLEER && !quirks) {
            vanleer_pair(u0[k] - um[k], up[k] - u0[k], sl, sh);
        } else {
            sl = slope<LIM>(u0[k] - um[k], up[k], u0[k], quirks);
            sh = slope<LIM>(up[k] - u0[k], u0[k], um[k], quirks);
        }
        lo[k] = u0[k] + T(0.5) * sl;
        hi[k] = u0[k] - T(0.5) * sh;
    }
}

// Face states from the 4 cells around face i+1/2 (MUSCL_Scheme.cpp:83-85, :99-101):
// U_L = U_i + 0.5*lim(U_i - U_{i-1}, U_{i+1} - U_i),
// U_R = U_{i+1} - 0.5*lim(U_{i+2} - U_{i+1}, U_{i+1} - U_i).
template <int LIM, typename T> IB_HD void recon(const T* u0, const T* u1, const T* u2, const T* u3, T* UL, T* UR,
                                                 bool quirks) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        UL[k] = u1[k] + T(0.5) * slope<LIM>(u1[k] - u0[k], u2[k], u1[k], quirks);
        UR[k] = u2[k] - T(0.5) * slope<LIM>(u3[k] - u2[k], u2[k], u1[k], quirks);
    }
}

// Gas and mesh constants, with the reciprocals of the constant divisors.
template <typename T> struct GasT {
    T gam, gm1, gp1, c1;  // gamma, gamma-1, gamma+1, (gamma-1)/2/gamma
    Rcp<T> r_gam, r_gm1, r_dx, r_dy;
};

// primitive view of a conserved state (rho, u, v, p), with 1/rho kept for reuse
template <typename T> struct PState {
    T rho, u, v, p, E;
    Rcp<T> ir;
};
// Conserve2Flux.cpp:66-69: u = rho u / rho, v = rho v / rho, p = (gamma-1)(E - rho(u^2+v^2)/2)
template <typename T> IB_HD PState<T> pstate(T gm1, const T* U) {
    PState<T> s;
    s.rho = U[0];
    s.ir = rcp_of(s.rho);
    s.u = dq(U[1], s.ir);
    s.v = dq(U[2], s.ir);
    s.p = gm1 * (U[3] - T(0.5) * s.rho * (s.u * s.u + s.v * s.v));
    s.E = U[3];
    return s;
}
// sqrt(gamma p / rho), where the reference calls pow(gamma p / rho, 0.5)
template <typename T> IB_HD T sound(const GasT<T>& g, const PState<T>& s) { return ib_sqrt(dq(g.gam * s.p, s.ir)); }

// Conserve2Flux.cpp:71-74 (x) / :106-109 (y): physical flux of a conserved state.
// The tangential momentum term is rho*u*v evaluated left to right in both directions.
template <bool Y, typename T> IB_HD void phys_flux(T rho, T u, T v, T p, T E, T* F) {
    if (!Y) { F[0] = rho * u; F[1] = rho * u * u + p; F[2] = rho * u * v; F[3] = (E + p) * u; }
    else { F[0] = rho * v; F[1] = rho * u * v; F[2] = rho * v * v + p; F[3] = (E + p) * v; }
}
template <bool Y, typename T> IB_HD void phys_flux(const PState<T>& s, T* F) { phys_flux<Y>(s.rho, s.u, s.v, s.p, s.E, F); }
template <typename T> IB_HD void primitives(T gm1, const T* U, T& rho, T& u, T& v, T& p) {
    rho = U[0]; u = U[1] / rho; v = U[2] / rho;
    p = gm1 * (U[3] - T(0.5) * rho * (u * u + v * v));
}
template <bool Y, typename T> IB_HD void flux_of(T gm1, const T* U, T* F) {
    PState<T> s = pstate(gm1, U);
    phys_flux<Y>(s, F);
}

// Corrected mode only (no reference counterpart): positivity of a state, the
// per-face first-order fallback and the density/pressure floors (applied to
// updated fluid cells: rho <- max(rho, RHO_FLOOR); if p <= P_FLOOR,
// E <- P_FLOOR/(gamma-1) + kinetic energy).
template <typename T> IB_HD bool physical(const PState<T>& s) { return s.rho > T(0) && s.p > T(0); }
constexpr double RHO_FLOOR = 1e-8, P_FLOOR = 1e-8;

// HLL/Riemann_solver.cpp:131-168 (x), :202-235 (y). Returns true when the
// first branch (S_lo >= 0) is taken; F is then F(U_lo) (the y solver's caller
// decides what that branch stores).
// (gamma+1) p* / 2 / gamma / p_K + (gamma-1)/2/gamma, in the reference's order
template <typename T> IB_HD T shock_factor_arg(const GasT<T>& g, T p_star, T pK) {
    T t = g.gp1 * p_star * T(0.5);
    t = dq(t, g.r_gam);
    t = ddiv(t, pK);
    return t + g.c1;
}
template <bool Y, typename T>
IB_HD bool hll_face(const GasT<T>& g, const PState<T>& L, const PState<T>& R, const T* Ulo, const T* Uhi, T* F,
                    bool quirks) {
    T unl = Y ? L.v : L.u, unh = Y ? R.v : R.u;
    T al = sound(g, L), ah = sound(g, R);
    T rho_bar = T(0.5) * (L.rho + R.rho);
    T a_bar = T(0.5) * (al + ah);
    T p_star = T(0.5) * (L.p + R.p) + T(0.5) * (unl - unh) * (rho_bar * a_bar);
    T Sl, Sh;
    if (quirks) {
        Sl = unl - al * ib_sqrt(shock_factor_arg(g, p_star, L.p));
        Sh = unh + ah * ib_sqrt(shock_factor_arg(g, p_star, R.p));
    } else {
        // corrected estimate (Toro, Riemann Solvers, 3rd ed., eq. 10.59-10.60): p* >= 0 and
        // the shock factor only where p* exceeds the side's pressure
        p_star = p_star > T(0) ? p_star : T(0);
        T ql = p_star <= L.p ? T(1) : ib_sqrt(shock_factor_arg(g, p_star, L.p));
        T qh = p_star <= R.p ? T(1) : ib_sqrt(shock_factor_arg(g, p_star, R.p));
        Sl = unl - al * ql;
        Sh = unh + ah * qh;
    }
    T Fl[4], Fh[4];
    phys_flux<Y>(L, Fl);
    if (Sl >= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = Fl[k];
        return true;
    }
    phys_flux<Y>(R, Fh);
    if (Sh <= T(0)) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = Fh[k];
        return false;
    }
    Rcp<T> iw = rcp_of(Sh - Sl);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] = dq(Sh * Fl[k] - Sl * Fh[k] + Sh * Sl * (Uhi[k] - Ulo[k]), iw);
    return false;
}

// StegerWarming/Riemann_solver.cpp:184-234 (x), :274-325 (y): F+(U_lo) + F-(U_hi)
// with eigenvalue smoothing delta = 1e-6. un/ut = normal/tangential velocity.
template <bool Y, typename T> IB_HD void sw_half(const GasT<T>& g, T rho, T un, T ut, T a, T h, bool plus, T* out) {
    const T d2 = T(1e-6) * T(1e-6);
    T s1 = ib_sqrt(un * un + d2);
    T m = un - a, q = un + a;
    T s3 = ib_sqrt(m * m + d2), s4 = ib_sqrt(q * q + d2);
    T l1 = plus ? T(0.5) * (un + s1) : T(0.5) * (un - s1);
    T l3 = plus ? T(0.5) * (m + s3) : T(0.5) * (m - s3);
    T l4 = plus ? T(0.5) * (q + s4) : T(0.5) * (q - s4);
    T u2 = Y ? (ut * ut + un * un) : (un * un + ut * ut);
    T twogm1 = T(2) * g.gm1;
    // rho * (...) / 2 / gamma
    out[0] = dq(rho * (twogm1 * l1 + l3 + l4) * T(0.5), g.r_gam);
    out[1] = dq(rho * (twogm1 * un * l1 + (un - a) * l3 + (un + a) * l4) * T(0.5), g.r_gam);
    out[2] = dq(rho * (twogm1 * ut * l1 + l3 * ut + l4 * ut) * T(0.5), g.r_gam);
    out[3] = dq(rho * (g.gm1 * u2 * l1 + (h - a * un) * l3 + (h + a * un) * l4) * T(0.5), g.r_gam);
}
template <bool Y, typename T> IB_HD void sw_face(const GasT<T>& g, const PState<T>& L, const PState<T>& R, T* F) {
    T al = sound(g, L), ah = sound(g, R);
    // e = p / (gamma-1) / rho ; h = e + u*u/2 + v*v/2 + p/rho
    T el = dq(dq(L.p, g.r_gm1), L.ir), eh = dq(dq(R.p, g.r_gm1), R.ir);
    T hl = el + L.u * L.u * T(0.5) + L.v * L.v * T(0.5) + dq(L.p, L.ir);
    T hh = eh + R.u * R.u * T(0.5) + R.v * R.v * T(0.5) + dq(R.p, R.ir);
    T fp[4], fn[4];
    sw_half<Y>(g, L.rho, Y ? L.v : L.u, Y ? L.u : L.v, al, hl, true, fp);
    sw_half<Y>(g, R.rho, Y ? R.v : R.u, Y ? R.u : R.v, ah, hh, false, fn);
    if (!Y) {
        F[0] = fp[0] + fn[0]; F[1] = fp[1] + fn[1]; F[2] = fp[2] + fn[2]; F[3] = fp[3] + fn[3];
    } else {
        F[0] = fp[0] + fn[0]; F[1] = fp[2] + fn[2]; F[2] = fp[1] + fn[1]; F[3] = fp[3] + fn[3];
    }
}

// PVRS/Riemann_solver.cpp:308-435 (x), :463-589 (y). The y solver's stray
// store at :550 (energy flux written through the state index macro into another
// cell's record) is not reproduced: in the reference's parallel builds its
// target is raced; here that branch leaves the energy slot as the reference's
// own slot would be left (the aliased x-flux), reported through `stray`.
template <bool Y, typename T> IB_HD void pvrs_face(const GasT<T>& g, const T* Ulo, const T* Uhi, T* F, bool& stray) {
    T rl, ul, vl, pl, rh, uh, vh, ph;
    primitives(g.gm1, Ulo, rl, ul, vl, pl);
    primitives(g.gm1, Uhi, rh, uh, vh, ph);
    T unl = Y ? vl : ul, utl = Y ? ul : vl, unh = Y ? vh : uh, uth = Y ? uh : vh;
    T al = ib_sqrt(g.gam * pl / rl), ah = ib_sqrt(g.gam * ph / rh);
    T rho_bar = T(0.5) * (rl + rh);
    T a_bar = T(0.5) * (al + ah);
    T p_star = T(0.5) * (pl + ph) + T(0.5) * (unl - unh) * (rho_bar * a_bar);
    T u_star = T(0.5) * (unl + unh) + T(0.5) * (pl - ph) / (rho_bar * a_bar);
    T rl_star = rl + (p_star - pl) / al / al;
    T rh_star = rh + (p_star - ph) / ah / ah;
    T Sl = unl - al * ib_sqrt(g.gp1 * p_star / T(2) / g.gam / pl + g.c1);
    T Sh = unh + ah * ib_sqrt(g.gp1 * p_star / T(2) / g.gam / ph + g.c1);
    T S_Hl = unl - al;
    T S_Tl = u_star - pow(al * (p_star / ph), g.c1);
    T S_Hh = unh + ah;
    T S_Th = u_star + pow(ah * (p_star / ph), g.c1);
    T f[4];  // normal mass, normal momentum, tangential momentum, energy
    stray = false;
    auto side = [&](T rho, T un, T ut, T p, T E) {
        f[0] = rho * un; f[1] = rho * un * un + p; f[2] = Y ? rho * ut * un : rho * un * ut; f[3] = (E + p) * un;
    };
    auto star = [&](T rs, T ut) {
        f[0] = rs * u_star; f[1] = rs * u_star * u_star + p_star; f[2] = rs * u_star * ut;
        f[3] = (g.gam * p_star / g.gm1 + T(0.5) * rs * (u_star * u_star + ut * ut)) * u_star;
    };
    auto fan = [&](T rho, T un, T ut, T p, T as) {
        T base = T(2) / g.gp1 + g.gm1 * (rho) / g.gp1 / as;
        T rf = rho * pow(base, T(2) / g.gm1);
        T uf = T(2) * (as + g.gm1 * un / T(2)) / g.gp1;
        T pf = p * pow(base, T(2) * g.gam / g.gm1);
        f[0] = rf * uf; f[1] = rf * uf * uf + pf; f[2] = rf * uf * ut;
        f[3] = (g.gam * pf / g.gm1 + T(0.5) * rf * (uf * uf + ut * ut)) * uf;
    };
    if (u_star > T(0)) {
        if (p_star > pl) { if (Sl > T(0)) side(rl, unl, utl, pl, Ulo[3]); else star(rl_star, utl); }
        else if (S_Hl > T(0)) side(rl, unl, utl, pl, Ulo[3]);
        else if (S_Tl < T(0)) star(rl_star, utl);
        else fan(rl, unl, utl, pl, al);
    } else {
        if (p_star > ph) {
            if (Sh < T(0)) { side(rh, unh, uth, ph, Uhi[3]); stray = Y; }
            else star(rh_star, uth);
        } else if (S_Hh < T(0)) side(rh, unh, uth, ph, Uhi[3]);
        else if (S_Th > T(0)) star(rh_star, uth);
        else fan(rh, unh, uth, ph, ah);
    }
    if (!Y) { F[0] = f[0]; F[1] = f[1]; F[2] = f[2]; F[3] = f[3]; }
    else { F[0] = f[0]; F[1] = f[2]; F[2] = f[1]; F[3] = f[3]; }
}

}  // namespace ibtfo
