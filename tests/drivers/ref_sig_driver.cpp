// A driver written the way the reference's test drivers are (Test/Sod/Sod.cpp:
// read the .inp, allocate the host arrays, Psy_coord, Level_Set, Scheme_Index,
// Initialize, Boundary, then TimeAdvance (or Advance) / Boundary / Global_Boundary /
// ComputeDt / CaculateMass per step), linked against libibtfo_amd.so through
// the reference signatures of include/ibtfo_ref.h. The four ibtfo_ref_* calls
// stand where the reference has its OpenACC data directives.
//
// Usage: ref_sig_driver case.inp outdir
// The .inp names the mask PNG (mask_file) and the inflow state (inflow_rho,
// inflow_u, inflow_v, inflow_p), which this driver's Initialize hook spreads
// over the domain. Writes outdir/U.bin (final U_OLD), outdir/XY.bin (XYCOORD
// after Level_Set), outdir/log.bin (dt, now_t, mass per step) as raw doubles.
#include <cstdio>
#include <cstdlib>
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

// the driver's Initialize hook (the reference keeps it in Test/<Case>/Initialize.cpp)
static void Initialize(const char* inp, int N_x, int N_y, int ng, double gamma, double* U_OLD, double* U_NEW) {
    const double rho = num(inp, "inflow_rho", 1.4), u = num(inp, "inflow_u", 3.0), v = num(inp, "inflow_v", 0.0),
                 p = num(inp, "inflow_p", 1.0);
    const size_t cells = (size_t)(N_x + 2 * ng) * (N_y + 2 * ng);
    for (size_t q = 0; q < cells; q++) {
        double* U = U_OLD + q * 4;
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = p / (gamma - 1) + 0.5 * rho * (u * u + v * v);
        for (int k = 0; k < 4; k++) U_NEW[q * 4 + k] = U[k];
    }
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
    if (argc < 3) {
        fprintf(stderr, "usage: %s case.inp outdir\n", argv[0]);
        return 2;
    }
    char* filename = argv[1];
    const std::string out = argv[2];
    const int N_x = (int)num(filename, "N_x", 100), N_y = (int)num(filename, "N_y", 100);
    const int ng = (int)num(filename, "num_ghost_cell", 2), max_iter = (int)num(filename, "max_iter", 10);
    const double lo_x = num(filename, "lo_x", 0), lo_y = num(filename, "lo_y", 0);
    const double hi_x = num(filename, "hi_x", 1), hi_y = num(filename, "hi_y", 1);
    const double Psy_time = num(filename, "Psy_time", 1), CFL_number = num(filename, "CFL_number", 0.5);
    const double gamma = num(filename, "gamma", 1.4);
    const double Psy_L = hi_x - lo_x, Psy_H = hi_y - lo_y;
    const size_t cells = (size_t)(N_x + 2 * ng) * (N_y + 2 * ng);
    std::vector<double> U_OLD(cells * 4), U_NEW(cells * 4), XYCOORD(cells * 6), SCHEME_IDX(cells * 16), U_TMP(4);
    double dt = 0.0, now_t = 0.0, mass = 0.0;
    int iter = 0;

    // was: #pragma acc enter data create(...)
    if (ibtfo_ref_enter(filename, U_OLD.data(), U_NEW.data(), XYCOORD.data(), SCHEME_IDX.data())) {
        fprintf(stderr, "ibtfo_ref_enter: %s\n", ibtfo_last_error());
        return 1;
    }
    Psy_coord(lo_x, lo_y, Psy_L, Psy_H, N_x, N_y, ng, XYCOORD.data(), 1, 0);
    Level_Set(filename, N_x, N_y, ng, XYCOORD.data(), 1, 0);
    Scheme_Index(N_x, N_y, ng, XYCOORD.data(), SCHEME_IDX.data(), 1, 0);
    save(out + "/XY.bin", XYCOORD.data(), XYCOORD.size());
    Initialize(filename, N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data());
    // was: #pragma acc update device(U_OLD, U_NEW)
    ibtfo_ref_update_device(U_OLD.data());
    Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), XYCOORD.data(), SCHEME_IDX.data(), 1, 0);
    ComputeDt(Psy_L, Psy_H, N_x, N_y, ng, gamma, CFL_number, U_OLD.data(), XYCOORD.data(), &dt, 1, 0);
    // clip_last_step = 1: a driver that shortens the last step to end exactly at
    // Psy_time and advances its clock by the dt it stepped with
    const int clip = (int)num(filename, "clip_last_step", 0);
    // call_advance = 1: step through Advance, as SingleStep's TimeAdvance does
    const int call_advance = (int)num(filename, "call_advance", 0);
    std::vector<double> log;
    while (now_t < Psy_time && iter < max_iter) {
        double step_dt = dt;
        if (clip && now_t + step_dt > Psy_time) step_dt = Psy_time - now_t;
        if (call_advance)
            Advance(Psy_L, Psy_H, N_x, N_y, ng, gamma, step_dt, U_OLD.data(), U_TMP.data(), U_NEW.data(),
                    XYCOORD.data(), 1, 0);
        else
            TimeAdvance(Psy_L, Psy_H, N_x, N_y, ng, gamma, step_dt, U_OLD.data(), U_TMP.data(), U_NEW.data(),
                        XYCOORD.data(), SCHEME_IDX.data(), 1, 0);
        Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), XYCOORD.data(), SCHEME_IDX.data(), 1, 0);
        Global_Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), 1, 0);
        ComputeDt(Psy_L, Psy_H, N_x, N_y, ng, gamma, CFL_number, U_OLD.data(), XYCOORD.data(), &dt, 1, 0);
        CaculateMass(Psy_L, Psy_H, N_x, N_y, ng, U_OLD.data(), XYCOORD.data(), &mass, 1, 0);
        now_t = now_t + (clip ? step_dt : dt);
        iter++;
        log.push_back(dt);
        log.push_back(now_t);
        log.push_back(mass);
    }
    // was: #pragma acc update self(U_OLD) / exit data
    ibtfo_ref_update_self(U_OLD.data());
    save(out + "/U.bin", U_OLD.data(), U_OLD.size());
    save(out + "/log.bin", log.data(), log.size());
    ibtfo_ref_exit(U_OLD.data());
    printf("ref_sig_driver: %d steps, t = %.9g\n", iter, now_t);
    return 0;
}
