// A driver that calls the scheme's inner operators the way the reference's
// schemes do (MUSCL_Scheme.cpp:63,123; HLL/Riemann_solver.cpp:21-24,122-123,
// 193-194), through the reference signatures of include/ibtfo_ref.h, on a
// scratch record U_TMP that it fills itself on the host.
//
// Usage: ref_record_driver case.inp rec.bin outdir
// rec.bin holds M*N*W doubles, W = 28 (16 when the .inp selects Scheme = Lax,
// whose record has 4 slots: Lax/EQDefine.H:5-6). For each operator the driver
// starts from rec.bin and writes the whole record afterwards:
//   x.bin   Riemann_solverX                                 (not for Lax)
//   y.bin   Riemann_solverY                                 (not for Lax)
//   fx.bin  Conserve2FluxX, slot 5 -> slot 3 (U_L -> F_L); Lax: slot 0 -> slot 2
//   fy.bin  Conserve2FluxY, slot 6 -> slot 4 (U_U -> G_U); Lax: slot 0 -> slot 3
//   f.bin   Conserve2Flux, slot 5 -> slots 3 (x) and 0 (y); Lax: slot 0 -> slots 2 and 3
//           (Lax/Lax_Scheme.cpp:41)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ibtfo.h"
#include "ibtfo_ref.h"

static double num(const char* inp, const char* key, double def) {
    double v = def;
    if (ibtfo_inp_number(inp, key, &v) < 0) {
        fprintf(stderr, "cannot read %s\n", inp);
        exit(2);
    }
    return v;
}

static void save(const std::string& path, const double* a, size_t n) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || fwrite(a, sizeof(double), n, f) != n) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s case.inp rec.bin outdir\n", argv[0]);
        return 2;
    }
    char* filename = argv[1];
    const std::string out = argv[3];
    const int N_x = (int)num(filename, "N_x", 100), N_y = (int)num(filename, "N_y", 100);
    const int ng = (int)num(filename, "num_ghost_cell", 2);
    const double gamma = num(filename, "gamma", 1.4);
    const double Psy_L = num(filename, "hi_x", 1) - num(filename, "lo_x", 0);
    const double Psy_H = num(filename, "hi_y", 1) - num(filename, "lo_y", 0);
    char scheme[64] = {0};
    ibtfo_inp_string(filename, "Scheme", scheme, (int)sizeof scheme);
    const bool lax = std::string(scheme).rfind("Lax", 0) == 0;
    const size_t W = lax ? 16 : 28;
    const size_t cells = (size_t)(N_x + 2 * ng) * (N_y + 2 * ng);
    std::vector<double> U_OLD(cells * 4), U_NEW(cells * 4), XYCOORD(cells * 6), SCHEME_IDX(cells * 16);
    std::vector<double> rec(cells * W), U_TMP(cells * W);
    FILE* f = fopen(argv[2], "rb");
    if (!f || fread(rec.data(), sizeof(double), rec.size(), f) != rec.size()) {
        fprintf(stderr, "cannot read %s (%zu doubles)\n", argv[2], rec.size());
        return 2;
    }
    fclose(f);

    // was: #pragma acc enter data create(...)
    if (ibtfo_ref_enter(filename, U_OLD.data(), U_NEW.data(), XYCOORD.data(), SCHEME_IDX.data())) {
        fprintf(stderr, "ibtfo_ref_enter: %s\n", ibtfo_last_error());
        return 1;
    }
    double* T = U_TMP.data();
    if (!lax) {
        U_TMP = rec;
        Riemann_solverX(Psy_L, Psy_H, N_x, N_y, ng, gamma, U_OLD.data(), T, 1, 0);
        save(out + "/x.bin", T, U_TMP.size());
        U_TMP = rec;
        Riemann_solverY(Psy_L, Psy_H, N_x, N_y, ng, gamma, U_OLD.data(), T, 1, 0);
        save(out + "/y.bin", T, U_TMP.size());
    }
    const int su = lax ? 0 : 5, sfx = lax ? 2 : 3, suy = lax ? 0 : 6, sgy = lax ? 3 : 4;
    U_TMP = rec;
    Conserve2FluxX(N_x, N_y, ng, gamma, &T[su * 4], &T[sfx * 4], 1, 0);
    save(out + "/fx.bin", T, U_TMP.size());
    U_TMP = rec;
    Conserve2FluxY(N_x, N_y, ng, gamma, &T[suy * 4], &T[sgy * 4], 1, 0);
    save(out + "/fy.bin", T, U_TMP.size());
    U_TMP = rec;
    if (lax) Conserve2Flux(N_x, N_y, ng, gamma, &T[0], &T[2 * 4], &T[3 * 4], 1, 0);
    else Conserve2Flux(N_x, N_y, ng, gamma, &T[5 * 4], &T[3 * 4], &T[0], 1, 0);
    save(out + "/f.bin", T, U_TMP.size());
    ibtfo_ref_exit(U_OLD.data());
    printf("ref_record_driver: %zu cells\n", cells);
    return 0;
}
