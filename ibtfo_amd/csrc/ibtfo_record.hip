// The reference's inner operators on its per-cell scratch record, for drivers
// that assemble their own scheme from them:
//   Riemann_solverX / Riemann_solverY   Src/Base/Euler/Riemann/Riemann_solver.H:22-42
//   Conserve2Flux / Conserve2FluxX / Conserve2FluxY   Src/Base/Common/Conserve2Flux.H:11-37
//
// The record (the reference's U_TMP) holds 7 slots of 4 doubles per cell,
// cell (i, j) at rec[(M*j + i)*28 + 4*slot + k] (Src/Base/Euler/Advection/
// MUSCL/EQDefine.H:33-44): slot 2 = F_OLD / G_OLD, 3 = F_L / G_D, 4 = F_R / G_U,
// 5 = U_L / U_D, 6 = U_R / U_U. The fused step kernel never materialises it;
// these operators exist so that the reference's scheme-level call sequence
// still links. One thread per cell; the per-face arithmetic is the step
// kernel's own (ibtfo_math.h), so a record face gives the same bits as the
// same face inside a fused step.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "ibtfo_ctx.h"

namespace ibtfo {
namespace {

constexpr int REC = 28;  // doubles per cell record
enum { SLOT_OLD = 2, SLOT_FLO = 3, SLOT_FHI = 4, SLOT_ULO = 5, SLOT_UHI = 6 };
constexpr int RECORD_BLOCK = 256;

struct RecordArgs {
    double* rec;
    long long ncells;
    GasT<double> g;
    int quirks;
};

GasT<double> gas_of(double gamma) {
    GasT<double> g;
    g.gam = gamma; g.gm1 = gamma - 1; g.gp1 = gamma + 1; g.c1 = (gamma - 1) / 2 / gamma;
    g.r_gam = Rcp<double>{g.gam, 1.0 / g.gam};
    g.r_gm1 = Rcp<double>{g.gm1, 1.0 / g.gm1};
    g.r_dx = g.r_dy = Rcp<double>{1.0, 1.0};
    return g;
}

// One face per cell record: the states in slots 5 and 6, the flux into slot 2.
// HLL (HLL/Riemann_solver.cpp:122-123, :193-194) first keeps the physical flux
// of both states in slots 3 and 4. In reference mode the y solvers store what
// the reference stores: HLL nothing when S_D >= 0 (:224-226: the slot keeps
// the x flux it aliases), PVRS no energy component on the branch of its stray
// store (PVRS/Riemann_solver.cpp:550, whose target is another cell's record:
// dropped, as in the step kernel).
template <int RS, bool Y> __global__ __launch_bounds__(RECORD_BLOCK) void k_riemann_record(RecordArgs a) {
    const long long c = (long long)blockIdx.x * RECORD_BLOCK + threadIdx.x;
    if (c >= a.ncells) return;
    double* r = a.rec + c * REC;
    double lo[4], hi[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { lo[k] = r[4 * SLOT_ULO + k]; hi[k] = r[4 * SLOT_UHI + k]; }
    const PState<double> L = pstate(a.g.gm1, lo), R = pstate(a.g.gm1, hi);
    double F[4];
    bool up = false, stray = false;
    if (RS == RS_SW) {
        sw_face<Y>(a.g, L, R, F);
    } else if (RS == RS_PVRS) {
        pvrs_face<Y>(a.g, lo, hi, F, stray);
    } else {
        double Fl[4], Fh[4];
        phys_flux<Y>(L, Fl);
        phys_flux<Y>(R, Fh);
#pragma unroll
        for (int k = 0; k < 4; k++) { r[4 * SLOT_FLO + k] = Fl[k]; r[4 * SLOT_FHI + k] = Fh[k]; }
        up = hll_face<Y>(a.g, L, R, lo, hi, F, a.quirks != 0);
    }
    double* o = r + 4 * SLOT_OLD;
    if (Y && a.quirks && RS == RS_HLL && up) return;
    o[0] = F[0]; o[1] = F[1]; o[2] = F[2];
    if (!(Y && a.quirks && RS == RS_PVRS && stray)) o[3] = F[3];
}

// Conserve2Flux.cpp:59-111: the physical flux of the state at u + c*stride,
// x flux to f + c*stride, y flux to g + c*stride (either may be null). The
// state is read whole before anything is written, so f or g may be u itself.
__global__ __launch_bounds__(RECORD_BLOCK) void k_flux_record(const double* u, double* f, double* g,
                                                              long long ncells, long long stride, double gm1) {
    const long long c = (long long)blockIdx.x * RECORD_BLOCK + threadIdx.x;
    if (c >= ncells) return;
    double U[4];
#pragma unroll
    for (int k = 0; k < 4; k++) U[k] = u[c * stride + k];
    const PState<double> s = pstate(gm1, U);
    double F[4], G[4];
    phys_flux<false>(s, F);
    phys_flux<true>(s, G);
    if (f)
#pragma unroll
        for (int k = 0; k < 4; k++) f[c * stride + k] = F[k];
    if (g)
#pragma unroll
        for (int k = 0; k < 4; k++) g[c * stride + k] = G[k];
}

template <bool Y> hipError_t launch_riemann(int rs, const RecordArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)((a.ncells + RECORD_BLOCK - 1) / RECORD_BLOCK)), block(RECORD_BLOCK);
    switch (rs) {
    case RS_PVRS: k_riemann_record<RS_PVRS, Y><<<grid, block, 0, s>>>(a); break;
    case RS_SW: k_riemann_record<RS_SW, Y><<<grid, block, 0, s>>>(a); break;
    default: k_riemann_record<RS_HLL, Y><<<grid, block, 0, s>>>(a); break;  // HLLC is HLL in the reference
    }
    return hipGetLastError();
}

hipError_t launch_flux(double gamma, const double* u, double* f, double* g, long long ncells, int stride,
                       hipStream_t s) {
    if (ncells == 0) return hipSuccess;
    const dim3 grid((unsigned)((ncells + RECORD_BLOCK - 1) / RECORD_BLOCK)), block(RECORD_BLOCK);
    k_flux_record<<<grid, block, 0, s>>>(u, f, g, ncells, stride, gamma - 1);
    return hipGetLastError();
}

int riemann_on_device(ibtfo_ctx* x, int dir, double* rec) {
    RecordArgs a;
    a.rec = rec;
    a.ncells = (long long)x->L.M * x->L.N;
    a.g = gas_of(x->cfg.gamma);
    a.quirks = x->cfg.quirks;
    hipError_t e = dir ? launch_riemann<true>(x->cfg.riemann, a, x->stream)
                       : launch_riemann<false>(x->cfg.riemann, a, x->stream);
    if (e != hipSuccess) {
        set_error("Riemann solver on the scratch record: %s", hipGetErrorString(e));
        return IBTFO_E_HIP;
    }
    return IBTFO_OK;
}

// a record launch covers M*N cells in blocks of 256: the grid must fit in x
constexpr long long RECORD_MAX_CELLS = (long long)0x7fffffff * RECORD_BLOCK;

}  // namespace
}  // namespace ibtfo

using namespace ibtfo;

extern "C" int ibtfo_riemann_record(ibtfo_ctx* x, int dir, double* rec) {
    if (!x || !rec || (dir != 0 && dir != 1)) { set_error("ibtfo_riemann_record: bad argument"); return IBTFO_E_ARG; }
    return riemann_on_device(x, dir, rec);
}

extern "C" int ibtfo_riemann_record_host(ibtfo_ctx* x, int dir, double* rec) {
    if (!x || !rec || (dir != 0 && dir != 1)) {
        set_error("ibtfo_riemann_record_host: bad argument");
        return IBTFO_E_ARG;
    }
    const size_t bytes = (size_t)x->L.M * x->L.N * REC * sizeof(double);
    double* d = nullptr;
    if (hipMalloc((void**)&d, bytes) != hipSuccess) {
        set_error("ibtfo_riemann_record_host: out of device memory (%zu bytes)", bytes);
        return IBTFO_E_NOMEM;
    }
    int rc = IBTFO_OK;
    hipError_t e = hipMemcpyAsync(d, rec, bytes, hipMemcpyHostToDevice, x->stream);
    if (e == hipSuccess) {
        rc = riemann_on_device(x, dir, d);
        if (rc == IBTFO_OK) e = hipMemcpyAsync(rec, d, bytes, hipMemcpyDeviceToHost, x->stream);
    }
    const hipError_t es = hipStreamSynchronize(x->stream);  // nothing queued still uses d
    if (e == hipSuccess) e = es;
    (void)hipFree(d);
    if (rc == IBTFO_OK && e != hipSuccess) {
        set_error("ibtfo_riemann_record_host: %s", hipGetErrorString(e));
        rc = IBTFO_E_HIP;
    }
    return rc;
}

static int flux_args_ok(const char* who, const double* u, const double* f, const double* g, long long ncells,
                        int stride) {
    if (!u || (!f && !g) || ncells < 0 || ncells > RECORD_MAX_CELLS || stride < 4) {
        set_error("%s: bad argument", who);
        return 0;
    }
    return 1;
}

extern "C" int ibtfo_flux_record(double gamma, const double* u, double* f, double* g, long long ncells, int stride,
                                 void* stream) {
    if (!flux_args_ok("ibtfo_flux_record", u, f, g, ncells, stride)) return IBTFO_E_ARG;
    hipError_t e = launch_flux(gamma, u, f, g, ncells, stride, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("ibtfo_flux_record: %s", hipGetErrorString(e));
        return IBTFO_E_HIP;
    }
    return IBTFO_OK;
}

// Host records: the states are gathered into a packed array (4 doubles per
// cell), the fluxes come back packed and are scattered to f and g. All of u is
// read before f or g is written, so they may share one record array.
extern "C" int ibtfo_flux_record_host(double gamma, const double* u, double* f, double* g, long long ncells,
                                      int stride) {
    if (!flux_args_ok("ibtfo_flux_record_host", u, f, g, ncells, stride)) return IBTFO_E_ARG;
    if (ncells == 0) return IBTFO_OK;
    const size_t n4 = (size_t)ncells * 4;
    std::vector<double> h, hf, hg;
    try {
        h.resize(n4);
        hf.resize(f ? n4 : 0);
        hg.resize(g ? n4 : 0);
    } catch (const std::bad_alloc&) {
        set_error("ibtfo_flux_record_host: out of host memory");
        return IBTFO_E_NOMEM;
    }
    for (long long c = 0; c < ncells; c++)
        for (int k = 0; k < 4; k++) h[c * 4 + k] = u[c * stride + k];
    double *du = nullptr, *df = nullptr, *dg = nullptr;
    auto release = [&]() { (void)hipFree(du); (void)hipFree(df); (void)hipFree(dg); };
    if (hipMalloc((void**)&du, n4 * sizeof(double)) != hipSuccess ||
        (f && hipMalloc((void**)&df, n4 * sizeof(double)) != hipSuccess) ||
        (g && hipMalloc((void**)&dg, n4 * sizeof(double)) != hipSuccess)) {
        release();
        set_error("ibtfo_flux_record_host: out of device memory");
        return IBTFO_E_NOMEM;
    }
    hipError_t e = hipMemcpy(du, h.data(), n4 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_flux(gamma, du, df, dg, ncells, 4, nullptr);
    if (e == hipSuccess && f) e = hipMemcpy(hf.data(), df, n4 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && g) e = hipMemcpy(hg.data(), dg, n4 * sizeof(double), hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) {
        set_error("ibtfo_flux_record_host: %s", hipGetErrorString(e));
        return IBTFO_E_HIP;
    }
    for (long long c = 0; c < ncells; c++)
        for (int k = 0; k < 4; k++) {
            if (f) f[c * stride + k] = hf[c * 4 + k];
            if (g) g[c * stride + k] = hg[c * 4 + k];
        }
    return IBTFO_OK;
}
