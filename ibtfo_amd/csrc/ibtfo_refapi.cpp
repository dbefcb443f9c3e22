// Thin C++ shim with the reference's free-function signatures, so a driver
// written against the reference (Test/*/<Case>.cpp) keeps its call sequence.
// Each function finds the solve by the host array it is given (the OpenACC
// present table's role) and forwards to the C ABI; a failing call prints the
// library's message and aborts, since these signatures have no error channel.
//
// Mapping (reference header:line -> here):
//   Advance       Src/Base/Euler/Advection/Advance.H:14-26  -> as TimeAdvance (SingleStep
//                 only: SingleStep's TimeAdvance is this call)
//   TimeAdvance   Src/Base/TimeAdvance/TimeAdvance.H:13-26  -> ibtfo_step_dt(ctx, dt): one
//                 step with the driver's dt; the fused step also applies Boundary,
//                 Global_Boundary and ComputeDt
//   Boundary      Src/BoundaryCondition/Boundary.H:11-20    -> before the first step:
//                 ibtfo_boundary; after a TimeAdvance: already applied
//   Global_Boundary Src/BoundaryCondition/Boundary.H:22-29  -> applied by the step
//   ComputeDt     Src/Base/Common/ComputeDt.H:12-23         -> dt kept by the step
//   CaculateMass  Src/Utils/CaculateMass.H:14-23            -> mass kept by the step
//   WriteData     Src/Utils/WriteData.H:14-27               -> state copied to the
//                 host U_OLD (the reference's update self), then the same files
//   Scheme_Index  Src/BoundaryCondition/Scheme_Index.H:13-19 -> built with the geometry
//   Riemann_solverX/Y Src/Base/Euler/Riemann/Riemann_solver.H:22-42 -> ibtfo_riemann_record_host
//                 on the driver's U_TMP (the solve's Riemann slot)
//   Conserve2Flux{,X,Y} Src/Base/Common/Conserve2Flux.H:11-37 -> ibtfo_flux_record_host
//   Level_Set     Src/Geometry/Level_Set.H:12-18             -> the driver's own
//                 Level_Set_function hook fills phi on the host; classification,
//                 normals and ghost records run on the device. With a `mask_file`
//                 key in the .inp, the mask level set of Test/PNGGeo/Level_Set.cpp
//                 (which reads Demo.png) runs on the device from that PNG instead
// `ndevices` must be 1 and `device` 0: the whole domain lives on one GPU.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "../../include/ibtfo.h"
#include "../../include/ibtfo_ref.h"

namespace {
struct Bound {
    ibtfo_ctx* ctx;
    ibtfo_config cfg;
    double *U_OLD, *U_NEW, *XY, *SCH;
    bool stepped;  // a TimeAdvance has run since the last host-side change
};
std::map<const double*, Bound*> g_present;

Bound* find(const double* p, const char* who) {
    auto it = g_present.find(p);
    if (it == g_present.end()) {
        fprintf(stderr, "ibtfo: %s: array not registered with ibtfo_ref_enter\n", who);
        abort();
    }
    return it->second;
}
void check(int rc, const char* who) {
    if (rc != IBTFO_OK) {
        fprintf(stderr, "ibtfo: %s failed (%d): %s\n", who, rc, ibtfo_last_error());
        abort();
    }
}
void one_device(int ndevices, int device, const char* who) {
    if (ndevices != 1 || device != 0) {
        fprintf(stderr, "ibtfo: %s: ndevices=%d device=%d; the domain is resident on one GPU\n", who, ndevices,
                device);
        abort();
    }
}
}  // namespace

// the driver's level-set hook (Src/Geometry/Level_Set_function.H:13-19), linked from its test directory
void Level_Set_function(char* filename, const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD,
                        const int ndevices, const int device) __attribute__((weak));

extern "C" {

// replaces `acc enter data create(...)`: one solve per .inp file and host array set
int ibtfo_ref_enter(const char* inp, double* U_OLD, double* U_NEW, double* XYCOORD, double* SCHEME_IDX) {
    Bound* b = new Bound();
    int rc = ibtfo_config_from_inp(inp, &b->cfg);
    if (rc) { delete b; return rc; }
    if ((rc = ibtfo_create(&b->cfg, &b->ctx))) { delete b; return rc; }
    b->U_OLD = U_OLD; b->U_NEW = U_NEW; b->XY = XYCOORD; b->SCH = SCHEME_IDX;
    for (const double* p : {(const double*)U_OLD, (const double*)U_NEW, (const double*)XYCOORD, (const double*)SCHEME_IDX})
        if (p) g_present[p] = b;
    return IBTFO_OK;
}
// replaces `acc update device(U_OLD, U_NEW)`
int ibtfo_ref_update_device(double* U_OLD) {
    Bound* b = find(U_OLD, "ibtfo_ref_update_device");
    b->stepped = false;
    return ibtfo_upload_state(b->ctx, b->U_OLD, b->U_NEW);
}
// replaces `acc update self(U_OLD)`
int ibtfo_ref_update_self(double* U_OLD) {
    Bound* b = find(U_OLD, "ibtfo_ref_update_self");
    return ibtfo_download_state(b->ctx, b->U_OLD);
}
// replaces `acc exit data`
int ibtfo_ref_exit(double* U_OLD) {
    Bound* b = find(U_OLD, "ibtfo_ref_exit");
    for (auto it = g_present.begin(); it != g_present.end();) it = (it->second == b) ? g_present.erase(it) : ++it;
    int rc = ibtfo_destroy(b->ctx);
    delete b;
    return rc;
}

}  // extern "C"

void Psy_coord(const double, const double, const double, const double, const int, const int, const int,
               double* XYCOORD, const int ndevices, const int device) {
    one_device(ndevices, device, "Psy_coord");
    Bound* b = find(XYCOORD, "Psy_coord");
    int M, N;
    ibtfo_dims(b->ctx, &M, &N);
    const double dx = (b->cfg.hi_x - b->cfg.lo_x) / b->cfg.N_x;
    const double Li = -dx * b->cfg.num_ghost_cell + b->cfg.lo_x + dx / 2;
    const double Hi = -dx * b->cfg.num_ghost_cell + b->cfg.lo_y + dx / 2;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            XYCOORD[((size_t)M * j + i) * 6] = Li + i * dx;
            XYCOORD[((size_t)M * j + i) * 6 + 1] = Hi + j * dx;
        }
}

void Level_Set(char* filename, const int N_x, const int N_y, const int ng, double* XYCOORD, const int ndevices,
               const int device) {
    one_device(ndevices, device, "Level_Set");
    Bound* b = find(XYCOORD, "Level_Set");
    char mask[1024];
    int n = ibtfo_inp_string(filename, "mask_file", mask, (int)sizeof mask);
    if (n > 0) {
        // ParamReader drops the blanks and keeps an inline comment: cut at '#'
        int k = 0;
        while (mask[k] && mask[k] != '#') k++;
        mask[k] = 0;
        check(ibtfo_geometry_png(b->ctx, mask), "Level_Set");
        check(ibtfo_download_geometry(b->ctx, XYCOORD), "Level_Set");
        return;
    }
    if (!Level_Set_function) {
        fprintf(stderr, "ibtfo: Level_Set: no Level_Set_function hook linked\n");
        abort();
    }
    Level_Set_function(filename, N_x, N_y, ng, XYCOORD, 1, 0);
    const size_t cells = (size_t)(N_x + 2 * ng) * (N_y + 2 * ng);
    std::vector<double> phi(cells);
    for (size_t q = 0; q < cells; q++) phi[q] = XYCOORD[q * 6 + 2];
    check(ibtfo_geometry(b->ctx, IBTFO_LS_FIELD, nullptr, 0, phi.data()), "Level_Set");
    check(ibtfo_download_geometry(b->ctx, XYCOORD), "Level_Set");
}

void Scheme_Index(const int, const int, const int, double* XYCOORD, double* SCH, const int ndevices, const int device) {
    one_device(ndevices, device, "Scheme_Index");
    Bound* b = find(XYCOORD, "Scheme_Index");
    if (SCH) check(ibtfo_download_scheme_index(b->ctx, SCH), "Scheme_Index");
}

void Boundary(const int, const int, const int, const double, double* U_OLD, double*, double*, double*,
              const int ndevices, const int device) {
    one_device(ndevices, device, "Boundary");
    Bound* b = find(U_OLD, "Boundary");
    if (b->stepped) return;  // applied by the fused step
    check(ibtfo_upload_state(b->ctx, b->U_OLD, b->U_NEW), "Boundary");
    check(ibtfo_boundary(b->ctx), "Boundary");
    check(ibtfo_download_state(b->ctx, b->U_OLD), "Boundary");
}

void Global_Boundary(const int, const int, const int, const double, double* U_OLD, double*, const int ndevices,
                     const int device) {
    one_device(ndevices, device, "Global_Boundary");
    (void)find(U_OLD, "Global_Boundary");  // applied by the fused step
}

static void ensure_prepared(Bound* b) {
    if (b->stepped) return;
    check(ibtfo_upload_state(b->ctx, b->U_OLD, b->U_NEW), "prepare");
    check(ibtfo_prepare(b->ctx, 0), "prepare");
}

void TimeAdvance(const double, const double, const int, const int, const int, const double, double dt,
                 double* U_OLD, double*, double*, double*, double*, const int ndevices, const int device) {
    one_device(ndevices, device, "TimeAdvance");
    Bound* b = find(U_OLD, "TimeAdvance");
    ensure_prepared(b);
    // SingleStep.cpp:5-20 advances by the dt it is given (a driver may have
    // changed the one ComputeDt returned, e.g. clipped it to Psy_time)
    check(ibtfo_step_dt(b->ctx, dt), "TimeAdvance");
    b->stepped = true;
}

// Advance.H:14-26: the scheme's update of U_NEW from U_OLD with the given dt, which
// is all SingleStep's TimeAdvance does (SingleStep.cpp:5-20), so the same fused
// step. Under RK2 the reference's Advance is one forward-Euler stage of Heun's
// step; the device runs whole steps, so that call is refused.
void Advance(const double, const double, const int, const int, const int, const double, double dt, double* U_OLD,
             double*, double*, double*, const int ndevices, const int device) {
    one_device(ndevices, device, "Advance");
    Bound* b = find(U_OLD, "Advance");
    if (b->cfg.time != IBTFO_SINGLESTEP) {
        fprintf(stderr, "ibtfo: Advance: the solve steps with RK2; call TimeAdvance (whole steps)\n");
        abort();
    }
    ensure_prepared(b);
    check(ibtfo_step_dt(b->ctx, dt), "Advance");
    b->stepped = true;
}

void ComputeDt(const double, const double, const int, const int, const int, const double, const double,
               double* U_OLD, double*, double* TMP_DT, const int ndevices, const int device) {
    one_device(ndevices, device, "ComputeDt");
    Bound* b = find(U_OLD, "ComputeDt");
    double dt = 0;
    if (b->stepped) check(ibtfo_status(b->ctx, nullptr, nullptr, &dt), "ComputeDt");
    else {
        ensure_prepared(b);
        check(ibtfo_compute_dt(b->ctx, &dt), "ComputeDt");
    }
    TMP_DT[device] = dt;
}

void CaculateMass(const double, const double, const int, const int, const int, double* U_OLD, double*,
                  double* MASS_DEVICE, const int ndevices, const int device) {
    one_device(ndevices, device, "CaculateMass");
    Bound* b = find(U_OLD, "CaculateMass");
    double m = 0;
    check(ibtfo_mass(b->ctx, &m), "CaculateMass");
    MASS_DEVICE[device] = m;
}

void WriteData(const double, const double, const double, const double, const int, const int, const int, int iter,
               double now_t, const double, double* U_OLD, double* XYCOORD, const int ndevices, const int device) {
    one_device(ndevices, device, "WriteData");
    Bound* b = find(U_OLD, "WriteData");
    if (b->stepped) check(ibtfo_download_state(b->ctx, b->U_OLD), "WriteData");
    check(ibtfo_vtk_write_host(&b->cfg, iter, now_t, U_OLD, XYCOORD, nullptr), "WriteData");
}

// The scheme's inner operators on the driver's scratch record U_TMP (MUSCL,
// MHM and Godunov call them with the whole record, all M*N cells). The solve
// is found by U_OLD; its grid and gamma must be the ones passed.
static void riemann_record(int dir, const int N_x, const int N_y, const int ng, const double gamma,
                           double* U_OLD, double* U_TMP, const int ndevices, const int device, const char* who) {
    one_device(ndevices, device, who);
    Bound* b = find(U_OLD, who);
    if (N_x != b->cfg.N_x || N_y != b->cfg.N_y || ng != b->cfg.num_ghost_cell || gamma != b->cfg.gamma) {
        fprintf(stderr, "ibtfo: %s: grid or gamma differs from the solve's .inp\n", who);
        abort();
    }
    if (b->cfg.scheme == IBTFO_LAX) {  // its record is 16 doubles a cell, and its build links no Riemann solver
        fprintf(stderr, "ibtfo: %s: the Lax scheme has no Riemann solver\n", who);
        abort();
    }
    check(ibtfo_riemann_record_host(b->ctx, dir, U_TMP), who);
}

void Riemann_solverX(const double, const double, const int N_x, const int N_y, const int ng, const double gamma,
                     double* U_OLD, double* U_TMP, const int ndevices, const int device) {
    riemann_record(0, N_x, N_y, ng, gamma, U_OLD, U_TMP, ndevices, device, "Riemann_solverX");
}

void Riemann_solverY(const double, const double, const int N_x, const int N_y, const int ng, const double gamma,
                     double* U_OLD, double* U_TMP, const int ndevices, const int device) {
    riemann_record(1, N_x, N_y, ng, gamma, U_OLD, U_TMP, ndevices, device, "Riemann_solverY");
}

// Conserve2Flux.cpp: M*N states at a stride of one scratch record, whose width
// the scheme's EQDefine.H fixes at link time: 4 slots of 4 doubles for Lax
// (Lax/EQDefine.H:5-6), 7 for the others (MUSCL/EQDefine.H:5-6). These calls
// carry no solve's array, so the width is the registered solves' scheme's.
static long long record_cells(int N_x, int N_y, int ng) { return (long long)(N_x + 2 * ng) * (N_y + 2 * ng); }
static int record_width(const char* who) {
    int w = 0;
    for (const auto& kv : g_present) {
        const int wk = kv.second->cfg.scheme == IBTFO_LAX ? 16 : 28;
        if (w && w != wk) {
            fprintf(stderr, "ibtfo: %s: the registered solves' schemes have scratch records of different widths\n", who);
            abort();
        }
        w = wk;
    }
    return w ? w : 28;
}

void Conserve2Flux(const int N_x, const int N_y, const int ng, const double gamma, double* U, double* F, double* G,
                   const int ndevices, const int device) {
    one_device(ndevices, device, "Conserve2Flux");
    check(ibtfo_flux_record_host(gamma, U, F, G, record_cells(N_x, N_y, ng), record_width("Conserve2Flux")),
          "Conserve2Flux");
}

void Conserve2FluxX(const int N_x, const int N_y, const int ng, const double gamma, double* U, double* F,
                    const int ndevices, const int device) {
    one_device(ndevices, device, "Conserve2FluxX");
    check(ibtfo_flux_record_host(gamma, U, F, nullptr, record_cells(N_x, N_y, ng), record_width("Conserve2FluxX")),
          "Conserve2FluxX");
}

void Conserve2FluxY(const int N_x, const int N_y, const int ng, const double gamma, double* U, double* G,
                    const int ndevices, const int device) {
    one_device(ndevices, device, "Conserve2FluxY");
    check(ibtfo_flux_record_host(gamma, U, nullptr, G, record_cells(N_x, N_y, ng), record_width("Conserve2FluxY")),
          "Conserve2FluxY");
}
