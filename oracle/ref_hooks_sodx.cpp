// Problem hook for the x-split Sod tube (diaphragm at x = x0, default 0.5), in the
// reference's hook signature (Src/Base/Common/Initialize.H:12-23) and AoS layout.
// The reference ships no such hook (its Sod initialiser splits on the row index,
// Test/Sod/Initialize.cpp:41). TEST INFRASTRUCTURE ONLY.
#include "ParamReader.H"
void Initialize(char* filename, const double, const double, const int N_x, const int N_y, const int ng,
                const double gamma, double* U_OLD, double* U_NEW, double* XY, const int, const int) {
    ParamReader rd;
    Params<double> para(rd.open(filename).numbers());
    const double s[8] = {para.get("rho_L", 1), para.get("u_L", 0), para.get("v_L", 0), para.get("p_L", 1),
                         para.get("rho_R", 0.125), para.get("u_R", 0), para.get("v_R", 0), para.get("p_R", 0.1)};
    const double x0 = para.get("x0", 0.5);
    const long cells = (long)(N_x + 2 * ng) * (N_y + 2 * ng);
    for (long q = 0; q < cells; q++) {
        const double* w = XY[q * 6] < x0 ? s : s + 4;
        U_OLD[q * 4 + 0] = w[0];
        U_OLD[q * 4 + 1] = w[0] * w[1];
        U_OLD[q * 4 + 2] = w[0] * w[2];
        U_OLD[q * 4 + 3] = w[3] / (gamma - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]);
        for (int k = 0; k < 4; k++) U_NEW[q * 4 + k] = U_OLD[q * 4 + k];
    }
}
