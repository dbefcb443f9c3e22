// Harness that drives the reference's own translation units, compiled from
// their sources under /root/reference by oracle/build_ref.py into
// oracle/_ref/libref_<combo>.so. TEST INFRASTRUCTURE ONLY: it exists so the
// oracle restatement can be checked bit-for-bit against the reference.
//
// The reference selects scheme/solver/limiter/boundary at link time
// (Test/Sod/Makefile:1-45), so each combination is a separate library exporting
// the same C entry points below. The call order reproduces the drivers:
// Test/Sod/Sod.cpp:94-181 and Test/InclineSod/InclineSod.cpp:73-165.
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ParamReader.H"

// Reference entry points (C++ linkage; signatures from the reference headers:
// Psy_coord.H, Level_Set.H, Scheme_Index.H, Initialize.H, Boundary.H,
// TimeAdvance.H, ComputeDt.H, CaculateMass.H).
void Psy_coord(const double, const double, const double, const double, const int, const int, const int,
               double*, const int, const int);
void Level_Set(char*, const int, const int, const int, double*, const int, const int);
void Scheme_Index(const int, const int, const int, double*, double*, const int, const int);
void Initialize(char*, const double, const double, const int, const int, const int, const double, double*,
                double*, double*, const int, const int);
void Boundary(const int, const int, const int, const double, double*, double*, double*, double*, const int,
              const int);
void Global_Boundary(const int, const int, const int, const double, double*, double*, const int, const int);
void TimeAdvance(const double, const double, const int, const int, const int, const double, double, double*,
                 double*, double*, double*, double*, const int, const int);
void ComputeDt(const double, const double, const int, const int, const int, const double, const double, double*,
               double*, double*, const int, const int);
void CaculateMass(const double, const double, const int, const int, const int, double*, double*, double*,
                  const int, const int);
void Riemann_solverX(const double, const double, const int, const int, const int, const double, double*, double*,
                     const int, const int);  // Riemann_solver.H:22-31
void Riemann_solverY(const double, const double, const int, const int, const int, const double, double*, double*,
                     const int, const int);  // Riemann_solver.H:33-42
void Conserve2Flux(const int, const int, const int, const double, double*, double*, double*, const int,
                   const int);  // Conserve2Flux.H:11-19
void Conserve2FluxX(const int, const int, const int, const double, double*, double*, const int, const int);
void Conserve2FluxY(const int, const int, const int, const double, double*, double*, const int, const int);

#ifndef REF_NUM_SCH
#define REF_NUM_SCH 1
#endif
#ifndef REF_TMP_W
#define REF_TMP_W 28
#endif
#ifndef REF_HAS_SCHEME_INDEX
#define REF_HAS_SCHEME_INDEX 0
#endif

extern "C" {

int ref_sch_width() { return REF_NUM_SCH; }
int ref_tmp_width() { return REF_TMP_W; }

// Runs a case described by an .inp file. Arrays are allocated here, zeroed, and
// copied out at the end (all may be null except where noted). Returns the number
// of steps taken (bounded by nsteps as well as by Psy_time / max_iter).
int ref_run(const char* inp, int nsteps, int init_boundary, double* U_OLD_out, double* U_NEW_out,
            double* XY_out, double* SCH_out, double* dt0_out, double* dt_log, double* t_log,
            double* mass_log) {
    ParamReader rd;
    Params<double> para(rd.open(inp).numbers());
    const int max_iter = int(para.get("max_iter", 100000));
    const double lo_x = para.get("lo_x", 0), lo_y = para.get("lo_y", 0);
    const double hi_x = para.get("hi_x", 1), hi_y = para.get("hi_y", 1);
    const int N_x = int(para.get("N_x", 500)), N_y = int(para.get("N_y", 200));
    const double Psy_time = para.get("Psy_time", 1.0);
    const double CFL_number = para.get("CFL_number", 0.3);
    const int ng = int(para.get("num_ghost_cell", 2));
    const double gamma = para.get("gamma", 1.4);
    const double Psy_L = hi_x - lo_x, Psy_H = hi_y - lo_y;
    const size_t cells = (size_t)(N_x + 2 * ng) * (N_y + 2 * ng);
    std::vector<double> U_OLD(cells * 4, 0.0), U_NEW(cells * 4, 0.0), XY(cells * 6, 0.0);
    std::vector<double> U_TMP(cells * REF_TMP_W, 0.0), SCH(cells * REF_NUM_SCH, 0.0);
    std::vector<char> fname(inp, inp + strlen(inp) + 1);
    double TMP_DT[1], MASS[1];

    Psy_coord(lo_x, lo_y, Psy_L, Psy_H, N_x, N_y, ng, XY.data(), 1, 0);
    Level_Set(fname.data(), N_x, N_y, ng, XY.data(), 1, 0);
#if REF_HAS_SCHEME_INDEX
    Scheme_Index(N_x, N_y, ng, XY.data(), SCH.data(), 1, 0);
#endif
    Initialize(fname.data(), Psy_L, Psy_H, N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), XY.data(), 1, 0);
    if (init_boundary) Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), XY.data(), SCH.data(), 1, 0);
    ComputeDt(Psy_L, Psy_H, N_x, N_y, ng, gamma, CFL_number, U_OLD.data(), XY.data(), TMP_DT, 1, 0);
    double dt = TMP_DT[0];
    if (dt0_out) *dt0_out = dt;
    double now_t = 0;
    int iter = 0;
    while (now_t < Psy_time && iter < max_iter && iter < nsteps) {
        TimeAdvance(Psy_L, Psy_H, N_x, N_y, ng, gamma, dt, U_OLD.data(), U_TMP.data(), U_NEW.data(), XY.data(),
                    SCH.data(), 1, 0);
        Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), XY.data(), SCH.data(), 1, 0);
        Global_Boundary(N_x, N_y, ng, gamma, U_OLD.data(), U_NEW.data(), 1, 0);
        ComputeDt(Psy_L, Psy_H, N_x, N_y, ng, gamma, CFL_number, U_OLD.data(), XY.data(), TMP_DT, 1, 0);
        CaculateMass(Psy_L, Psy_H, N_x, N_y, ng, U_OLD.data(), XY.data(), MASS, 1, 0);
        dt = TMP_DT[0];
        now_t = now_t + dt;
        if (dt_log) dt_log[iter] = dt;
        if (t_log) t_log[iter] = now_t;
        if (mass_log) mass_log[iter] = MASS[0];
        iter++;
    }
    if (U_OLD_out) memcpy(U_OLD_out, U_OLD.data(), U_OLD.size() * sizeof(double));
    if (U_NEW_out) memcpy(U_NEW_out, U_NEW.data(), U_NEW.size() * sizeof(double));
    if (XY_out) memcpy(XY_out, XY.data(), XY.size() * sizeof(double));
    if (SCH_out) memcpy(SCH_out, SCH.data(), SCH.size() * sizeof(double));
    return iter;
}

// One TimeAdvance (= Advance) from a caller-provided state, for unit-level checks.
// U_TMP must hold ref_tmp_width() doubles per cell and persists across calls.
void ref_advance(double L, double H, int nx, int ny, int ng, double gamma, double dt, double* U_OLD, double* U_TMP,
                 double* U_NEW, double* XY) {
    TimeAdvance(L, H, nx, ny, ng, gamma, dt, U_OLD, U_TMP, U_NEW, XY, nullptr, 1, 0);
}

// The linked Riemann solver's Riemann_solverX (ydir 0) / Riemann_solverY (1) on
// a whole scratch record (M*N cells x 28 doubles); U_OLD is only named in its
// data clause. Lax libraries link no Riemann solver.
#if REF_TMP_W == 28
void ref_riemann_record(int ydir, int nx, int ny, int ng, double gamma, double* U_OLD, double* U_TMP) {
    if (ydir) Riemann_solverY(1.0, 1.0, nx, ny, ng, gamma, U_OLD, U_TMP, 1, 0);
    else Riemann_solverX(1.0, 1.0, nx, ny, ng, gamma, U_OLD, U_TMP, 1, 0);
}
#endif

// Conserve2Flux (which 2), Conserve2FluxX (0) or Conserve2FluxY (1) on M*N
// states at the record stride (28 doubles)
void ref_conserve2flux(int which, int nx, int ny, int ng, double gamma, double* U, double* F, double* G) {
    if (which == 2) Conserve2Flux(nx, ny, ng, gamma, U, F, G, 1, 0);
    else if (which == 1) Conserve2FluxY(nx, ny, ng, gamma, U, G, 1, 0);
    else Conserve2FluxX(nx, ny, ng, gamma, U, F, 1, 0);
}

// Parsed .inp numbers()/strings() maps, written as key 0x1f value 0x1e records into buf.
int ref_parse(const char* inp, int strings, char* buf, int cap) {
    ParamReader rd;
    std::string out;
    rd.open(inp);
    if (strings) {
        std::map<std::string, std::string> m = rd.strings();
        for (auto& kv : m) out += kv.first + "\x1f" + kv.second + "\x1e";
    } else {
        std::map<std::string, double> m = rd.numbers();
        char tmp[64];
        for (auto& kv : m) {
            snprintf(tmp, sizeof tmp, "%.17g", kv.second);
            out += kv.first + "\x1f" + tmp + "\x1e";
        }
    }
    if ((int)out.size() + 1 > cap) return -(int)out.size() - 1;
    memcpy(buf, out.c_str(), out.size() + 1);
    return (int)out.size();
}

}  // extern "C"
