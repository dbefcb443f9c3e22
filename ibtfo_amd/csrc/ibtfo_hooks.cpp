// Initialize hooks (the reference's per-problem user code), evaluated on the
// host once per solve and uploaded. Transcendental calls stay on the host so the
// initial state uses the same libm the reference's CPU build uses.
#include <cmath>
#include <cstring>

#include "../../include/ibtfo.h"
#include "ibtfo_host.h"

namespace ibtfo {

static inline void to_cons(double gam, double rho, double u, double v, double p, double* U) {
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = p / (gam - 1) + 0.5 * rho * (u * u + v * v);
}

int host_initialize(int kind, const double* prm, int nprm, int nx, int ny, int ng, double gam, const double* xs,
                    const double* ys, const unsigned char* type, double* U) {
    static const int need[] = {8, 9, 2, 9, 4, 7, 6};
    if (kind < 0 || kind > 6) { set_error("unknown initialiser %d", kind); return IBTFO_E_ARG; }
    if (!prm || nprm < need[kind]) { set_error("initialiser %d needs %d parameters", kind, need[kind]); return IBTFO_E_ARG; }
    const int M = nx + 2 * ng, N = ny + 2 * ng;
    const double pi = M_PI;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            const size_t q = (size_t)M * j + i;
            double* u = U + q * 4;
            const double x = xs[q], y = ys[q];
            const int t = type[q];
            switch (kind) {
                case IBTFO_INIT_SOD_ROWS: {  // Test/Sod/Initialize.cpp:35-50: split on the row index
                    const double* s = (j < (nx + 2 * ng) / 2) ? prm : prm + 4;
                    to_cons(gam, s[0], s[1], s[2], s[3], u);
                    break;
                }
                case IBTFO_INIT_INCLINE: {  // Test/InclineSod/Initialize.cpp:38-72
                    double angle = 3.1415926535 * prm[0] / 180;
                    if (t == 0 || t == 2) {
                        const double* s =
                            (y < -(1 / tan(angle)) * x + 0.5 * (1 / tan(angle)) + 0.5) ? prm + 1 : prm + 5;
                        to_cons(gam, s[0], s[1], s[2], s[3], u);
                    } else {
                        u[0] = -1; u[1] = 0; u[2] = 0; u[3] = -1;
                    }
                    break;
                }
                case IBTFO_INIT_VORTEX: {  // Test/Vortex/Initialize.cpp:20-97
                    const double rho0 = 1.0, p0 = 2.0, xlc = 0.50, ylc = 0.50, rv = 0.40;
                    const double ua = prm[0] * rv * pi;
                    double ur = 0, pr = 0, uu = 0, vv = 0;
                    if (t == 0 || t == 2) {
                        double r = pow((x - xlc) * (x - xlc) + (y - ylc) * (y - ylc), 0.5);
                        double ph = atan((y - ylc) / (x - xlc));
                        if (x - xlc < 0.0) ph = ph + pi;
                        if (r < rv) {
                            if (r < rv / 2.0) {
                                ur = ua * 2.0 * r / rv;
                                pr = (r / rv) * (r / rv) + 1.0 - 2.0 * log(2.0);
                            } else {
                                ur = ua * 2.0 * (1.0 - r / rv);
                                pr = (r / rv) * (r / rv) + 3.0 - 4.0 * r / rv + 2.0 * log(r / rv);
                            }
                            // one sincos call, which is what g++ emits for the hook's
                            // sin(phi) and cos(phi) pair (results can differ by an ulp
                            // from separate sin and cos calls)
                            double sp, cp;
                            sincos(ph, &sp, &cp);
                            uu = -sp * ur;
                            vv = cp * ur;
                        }
                    }
                    u[0] = rho0;
                    u[1] = rho0 * uu;
                    u[2] = rho0 * vv;
                    u[3] = (p0 + 2.0 * rho0 * ua * ua * pr) / (gam - 1) + 0.5 * rho0 * (uu * uu + vv * vv);
                    break;
                }
                case IBTFO_INIT_SOD_X: {
                    const double* s = (x < prm[0]) ? prm + 1 : prm + 5;
                    to_cons(gam, s[0], s[1], s[2], s[3], u);
                    break;
                }
                case IBTFO_INIT_UNIFORM:
                    to_cons(gam, prm[0], prm[1], prm[2], prm[3], u);
                    break;
                case IBTFO_INIT_SHOCK: {  // Test/PNGGeo/Initialize.cpp:29-55, in its operation order
                    const double Ms = prm[1], rho_1 = prm[2], u_1 = prm[3], v_1 = prm[4], p_1 = prm[5];
                    if (x > prm[0]) {
                        to_cons(gam, rho_1, u_1, v_1, p_1, u);
                    } else {
                        double rho_2 = rho_1 * (gam + 1) * Ms * Ms / (2 + (gam - 1) * Ms * Ms);
                        double p_2 = p_1 * ((2 * gam) * Ms * Ms / (gam + 1) - (gam - 1) / (gam + 1));
                        double u_2 = Ms * (1 - rho_1 / rho_2);
                        u[0] = rho_2; u[1] = rho_2 * u_2; u[2] = 0;
                        u[3] = p_2 / (gam - 1) + 0.5 * rho_2 * (u_2 * u_2);
                    }
                    break;
                }
                default: {  // isentropic vortex: T = 1 - (g-1) b^2/(8 g pi^2) e^(1-r^2), rho = T^(1/(g-1)), p = rho T
                    double dxv = x - prm[0], dyv = y - prm[1], r2 = dxv * dxv + dyv * dyv, beta = prm[2];
                    double T = 1 - (gam - 1) * beta * beta / (8 * gam * pi * pi) * exp(1 - r2);
                    double rho = prm[3] * pow(T, 1 / (gam - 1));
                    double p = prm[6] / prm[3] * rho * T;
                    double f = beta / (2 * pi) * exp(0.5 * (1 - r2));
                    to_cons(gam, rho, prm[4] - f * dyv, prm[5] + f * dxv, p, u);
                }
            }
        }
    return IBTFO_OK;
}

}  // namespace ibtfo
