// C ABI of the IBTFO MI355X library: context, device memory, geometry, state
// exchange and the device-resident time loop (include/ibtfo.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ibtfo.h"
#include "ibtfo_geom.h"
#include "ibtfo_host.h"
#include "ibtfo_internal.h"
#include "ibtfo_ctx.h"

using namespace ibtfo;

namespace ibtfo {
static thread_local char g_err[512];
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
}  // namespace ibtfo


#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return IBTFO_E_HIP;                                                    \
        }                                                                          \
    } while (0)

static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

// Device memory of a context: from its arena when it belongs to a batch, else hipMalloc.
template <class P> static hipError_t dalloc(ibtfo_ctx* x, P** p, size_t bytes) {
    if (x->arena) {
        *p = (P*)x->arena->alloc(bytes);
        return *p ? hipSuccess : hipErrorOutOfMemory;
    }
    return hipMalloc((void**)p, bytes);
}
// A captured step graph holds the kernel arguments of its capture (buffer
// pointers, ghost / group counts, grid shape): anything that changes them
// drops the graph, and the next ibtfo_step_graph captures a new one.
static void drop_graph(ibtfo_ctx* x) {
    if (x->graph) {
        (void)hipStreamSynchronize(x->stream);  // no replay of it still in flight
        (void)hipGraphExecDestroy(x->graph);
    }
    x->graph = nullptr;
    x->graph_chunk = 0;
}

static void dfree(ibtfo_ctx* x, void* p) {
    if (!p) return;
    if (x->arena) x->arena->free(p);
    else (void)hipFree(p);
}
// Copies ordered on the context's stream (never the null stream, which would
// wait for every other blocking stream of the device).
static hipError_t scopy(ibtfo_ctx* x, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, x->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(x->stream);
}

static Layout make_layout(const ibtfo_config* c) {
    Layout L;
    L.nx = c->N_x; L.ny = c->N_y; L.ng = c->num_ghost_cell;
    L.M = L.nx + 2 * L.ng; L.N = L.ny + 2 * L.ng;
    L.off = (16 - L.ng % 16) % 16;
    L.pitch = (int)round_up(L.M + L.off, 16);
    L.plane = round_up((long long)L.off + (long long)L.pitch * L.N + 16, 64);
    return L;
}

// Step-kernel grid: strips of STRIP_W columns x `rows` rows. rows_override = 0
// picks the default shape (about 8192 blocks, 32 waves per CU to cover the
// latency of a 3-wave occupancy, and at least 4 rows per strip);
// ibtfo_set_rows / ibtfo_autotune choose another.
static int step_grid(const Layout& L, int rows_override, int* nbx, int* nby, int* rows) {
    *nbx = (L.nx + STRIP_W - 1) / STRIP_W;
    int want_y = std::max(1, 8192 / *nbx);
    int r = rows_override > 0 ? rows_override : std::max(4, (L.ny + want_y - 1) / want_y);
    r = std::min(r, std::max(L.ny, 1));
    *rows = r;
    *nby = (L.ny + r - 1) / r;
    return 0;
}

static int validate(const ibtfo_config* c) {
    if (c->num_ghost_cell < 2 || c->num_ghost_cell > 4) { set_error("num_ghost_cell must be 2..4"); return IBTFO_E_ARG; }
    if (c->N_x < 4 || c->N_y < 2 * c->num_ghost_cell) { set_error("grid too small"); return IBTFO_E_ARG; }
    if ((long long)c->N_x * c->N_y > (1ll << 31)) { set_error("grid too large"); return IBTFO_E_ARG; }
    if (c->scheme < 0 || c->scheme > 3 || c->riemann < 0 || c->riemann > 3 || c->limiter < 0 || c->limiter > 2 ||
        c->boundary < 0 || c->boundary > 7 || c->edge < 0 || c->edge > 3) {
        set_error("slot value out of range");
        return IBTFO_E_ARG;
    }
    if (c->time != IBTFO_SINGLESTEP && c->time != IBTFO_RK2) { set_error("unknown Time slot"); return IBTFO_E_ARG; }
    if (c->boundary == IBTFO_SIMPLE_BOUND && c->scheme == IBTFO_LAX && c->edge == IBTFO_EDGE_REFERENCE) {
        set_error("SimpleBound with Lax and the reference edge hook is not supported");
        return IBTFO_E_ARG;
    }
    if (c->boundary == IBTFO_CUTCELL && (c->quirks || c->time != IBTFO_SINGLESTEP || c->scheme == IBTFO_LAX)) {
        set_error("the cut-cell boundary needs quirks = 0, Time = SingleStep and a reconstructing scheme");
        return IBTFO_E_ARG;
    }
    if (c->time == IBTFO_RK2 && c->scheme == IBTFO_LAX) {
        set_error("Time=RK2 needs a reconstructing scheme (MUSCL, MHM, Godunov)");
        return IBTFO_E_ARG;
    }
    if (c->precision != IBTFO_FP64 && c->precision != IBTFO_MIXED) { set_error("unknown precision"); return IBTFO_E_ARG; }
    if (c->precision == IBTFO_MIXED && (c->quirks || c->scheme == IBTFO_LAX)) {
        set_error("mixed precision needs quirks = 0 and a reconstructing scheme (MUSCL, MHM, Godunov)");
        return IBTFO_E_ARG;
    }
    if (!(c->gamma > 1.0) || !(c->hi_x > c->lo_x) || !(c->hi_y > c->lo_y)) { set_error("bad physical parameters"); return IBTFO_E_ARG; }
    return IBTFO_OK;
}

extern "C" const char* ibtfo_last_error(void) { return g_err; }

extern "C" long long ibtfo_footprint(const ibtfo_config* c) {
    if (!c) return -1;
    Layout L = make_layout(c);
    long long b = 8 * L.plane * 8;            // two 4-plane state buffers
    b += L.plane;                              // types
    b += 3 * 8 * L.plane;                      // phi, normals
    b += (long long)sizeof(Ctl) + (1ll << 20);  // control, partials
    b += (long long)(2 * (L.M + L.N) * L.ng) * (sizeof(Ghost) + 32);  // ghost list upper bound
    return b;
}

void ibtfo::free_ctx(ibtfo_ctx* x) {
    if (!x) return;
    if (x->graph) (void)hipGraphExecDestroy(x->graph);
    for (void* p : {(void*)x->state[0], (void*)x->state[1], (void*)x->type, (void*)x->phi, (void*)x->gnx,
                    (void*)x->gny, (void*)x->ghosts, (void*)x->gvals, (void*)x->dcount, (void*)x->ctl,
                    (void*)x->mass_part, (void*)x->corner, (void*)x->cuts, (void*)x->cut_index, (void*)x->groups,
                    (void*)x->stype, (void*)x->merge_part, (void*)x->fold, (void*)x->unew0, (void*)x->df_part})
        dfree(x, p);
    if (x->own_stream) (void)hipStreamDestroy(x->own_stream);
    delete x;
}

namespace ibtfo {
// A context on `stream` (its own non-blocking stream when null) with device
// memory from `arena` (hipMalloc when null); rows = step-kernel strip height (0 =
// default). Everything the kernels may read before writing it is cleared.
int ctx_create(const ibtfo_config* cfg, Arena* arena, hipStream_t stream, int rows, ibtfo_ctx** out) {
    if (!cfg || !out) { set_error("null argument"); return IBTFO_E_ARG; }
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    ibtfo_ctx* x = new ibtfo_ctx();  // value-initialised: every plain member starts at zero
    x->cfg = *cfg;
    x->L = make_layout(cfg);
    x->arena = arena;
    x->rows_override = rows;
    const Layout& L = x->L;
    if (hipGetDevice(&x->device) != hipSuccess) { delete x; set_error("no HIP device"); return IBTFO_E_HIP; }
    auto fail = [&](const char* what) {
        set_error("allocation of %s failed (%s)", what, arena ? "batch arena full" : hipGetErrorString(hipGetLastError()));
        free_ctx(x);
        return IBTFO_E_NOMEM;
    };
    if (dalloc(x, &x->state[0], 4 * L.plane * sizeof(double)) != hipSuccess) return fail("state");
    if (dalloc(x, &x->state[1], 4 * L.plane * sizeof(double)) != hipSuccess) return fail("state");
    if (dalloc(x, &x->type, L.plane) != hipSuccess) return fail("types");
    if (dalloc(x, &x->phi, L.plane * sizeof(double)) != hipSuccess) return fail("phi");
    if (dalloc(x, &x->gnx, L.plane * sizeof(double)) != hipSuccess) return fail("normals");
    if (dalloc(x, &x->gny, L.plane * sizeof(double)) != hipSuccess) return fail("normals");
    if (dalloc(x, &x->ctl, sizeof(Ctl)) != hipSuccess) return fail("control");
    if (dalloc(x, &x->fold, FOLD_MAX * PART_STRIDE * sizeof(double)) != hipSuccess) return fail("partials");
    if (dalloc(x, &x->dcount, sizeof(unsigned int)) != hipSuccess) return fail("counter");
    int nbx, nby, r;
    step_grid(L, rows, &nbx, &nby, &r);
    x->mass_cap = arena ? nbx * nby : std::max(nbx * nby, 4096);
    if (dalloc(x, &x->mass_part, PART_STRIDE * (size_t)x->mass_cap * sizeof(double)) != hipSuccess)
        return fail("partials");
    if (dalloc(x, &x->corner, 2 * 2 * L.ng * 4 * sizeof(double)) != hipSuccess) return fail("corner");
    const size_t df_bytes = 2 * (size_t)DF_MAX_REC * PART_STRIDE * sizeof(double);
    if (!arena && dalloc(x, &x->df_part, df_bytes) != hipSuccess) return fail("partials");
    if (stream) {
        x->stream = stream;
    } else {
        if (hipStreamCreateWithFlags(&x->own_stream, hipStreamNonBlocking) != hipSuccess) return fail("stream");
        x->stream = x->own_stream;
    }
    if (hipMemsetAsync(x->state[0], 0, 4 * L.plane * sizeof(double), x->stream) != hipSuccess ||
        hipMemsetAsync(x->state[1], 0, 4 * L.plane * sizeof(double), x->stream) != hipSuccess ||
        hipMemsetAsync(x->type, 2, L.plane, x->stream) != hipSuccess ||
        hipMemsetAsync(x->ctl, 0, sizeof(Ctl), x->stream) != hipSuccess ||
        hipMemsetAsync(x->phi, 0, L.plane * sizeof(double), x->stream) != hipSuccess ||
        hipMemsetAsync(x->gnx, 0, L.plane * sizeof(double), x->stream) != hipSuccess ||
        hipMemsetAsync(x->gny, 0, L.plane * sizeof(double), x->stream) != hipSuccess ||
        hipMemsetAsync(x->corner, 0, 2 * 2 * L.ng * 4 * sizeof(double), x->stream) != hipSuccess ||
        (x->df_part && hipMemsetAsync(x->df_part, 0, df_bytes, x->stream) != hipSuccess))
        return fail("initial clear");  // stream-ordered before everything else on the solve's stream
    const double Psy_L = cfg->hi_x - cfg->lo_x;
    x->dxc = Psy_L / cfg->N_x;  // Psy_coord.cpp:21-22 (y spaced with dx as well)
    x->Li = -x->dxc * L.ng + cfg->lo_x + x->dxc / 2;
    x->Hi = -x->dxc * L.ng + cfg->lo_y + x->dxc / 2;
    *out = x;
    return IBTFO_OK;
}
}  // namespace ibtfo

extern "C" int ibtfo_create(const ibtfo_config* cfg, ibtfo_ctx** out) {
    return ctx_create(cfg, nullptr, nullptr, 0, out);
}

extern "C" int ibtfo_destroy(ibtfo_ctx* x) {
    if (!x) return IBTFO_E_ARG;
    (void)hipStreamSynchronize(x->stream);
    free_ctx(x);
    return IBTFO_OK;
}

extern "C" int ibtfo_set_stream(ibtfo_ctx* x, void* s) {
    if (!x) return IBTFO_E_ARG;
    HIPCHK(hipStreamSynchronize(x->stream));  // work queued on the old stream ends before the new one starts
    x->stream = s ? (hipStream_t)s : x->own_stream;
    drop_graph(x);
    return IBTFO_OK;
}

extern "C" int ibtfo_dims(const ibtfo_ctx* x, int* M, int* N) {
    if (!x) return IBTFO_E_ARG;
    if (M) *M = x->L.M;
    if (N) *N = x->L.N;
    return IBTFO_OK;
}

extern "C" int ibtfo_sch_width(const ibtfo_ctx* x) {
    switch (x->cfg.boundary) {
        case IBTFO_IDW_GFM: case IBTFO_CUTCELL: return 8;
        case IBTFO_LINEAR_GFM: case IBTFO_RIEMANN_GFM: return 4;
        case IBTFO_BCGFM: case IBTFO_HBCGFM: return 16;
        case IBTFO_SIMPLE_BOUND: return 2;
        default: return 1;
    }
}
extern "C" int ibtfo_num_ghosts(const ibtfo_ctx* x) { return x ? x->nghost : -1; }
extern "C" int ibtfo_write_vtk(ibtfo_ctx* x, const char* dir, int iter, double now_t) {
    if (!x) { set_error("null context"); return IBTFO_E_ARG; }
    if (!x->have_geometry || !x->have_state) { set_error("ibtfo_write_vtk needs geometry and state"); return IBTFO_E_STATE; }
    const size_t cells = (size_t)x->L.M * x->L.N;
    std::vector<double> U(cells * 4), XY(cells * 6);
    int rc = ibtfo_download_state(x, U.data());
    if (!rc) rc = ibtfo_download_geometry(x, XY.data());
    if (!rc) rc = write_vtk(&x->cfg, iter, now_t, U.data(), XY.data(), dir);
    return rc;
}

extern "C" int ibtfo_cut_counts(const ibtfo_ctx* x, int* cut_cells, int* merge_groups) {
    if (!x) { set_error("null context"); return IBTFO_E_ARG; }
    if (cut_cells) *cut_cells = x->ncut;
    if (merge_groups) *merge_groups = x->ngroups;
    return IBTFO_OK;
}

// --------------------------------------------------------------- geometry
// Runs the geometry kernels for a prepared GeomArgs (hook constants and inputs
// set) and builds the ghost list; frees `dphi` (a device input of the hook).
// Cut-cell geometry after geom_run: dense scratch, records, types, merge
// targets, ghosts; then the compact records, the index plane and the groups.
static int cut_geometry(ibtfo_ctx* x, GeomArgs& g) {
    const Layout& L = x->L;
    dfree(x, x->cuts); x->cuts = nullptr;
    dfree(x, x->cut_index); x->cut_index = nullptr;
    dfree(x, x->groups); x->groups = nullptr;
    x->ncut = x->ngroups = 0;
    double* dense = nullptr;
    if (dalloc(x, &dense, (size_t)L.M * L.N * 8 * sizeof(double)) != hipSuccess ||
        dalloc(x, &x->cut_index, (size_t)L.plane * sizeof(int)) != hipSuccess) {
        dfree(x, dense);
        set_error("cut-cell geometry: out of device memory");
        return IBTFO_E_NOMEM;
    }
    g.dcut = dense;
    g.cut_index = x->cut_index;
    g.ccount = x->dcount;
    unsigned int zero[2] = {0, 0}, cnt[2] = {0, 0};
    unsigned int* gcount = nullptr;
    if (dalloc(x, &gcount, sizeof(unsigned int)) != hipSuccess) { dfree(x, dense); return IBTFO_E_NOMEM; }
    g.gcount = gcount;
    auto pass = [&](int fill) -> hipError_t {
        hipError_t e = hipMemcpyAsync(x->dcount, &zero[0], sizeof(unsigned), hipMemcpyHostToDevice, x->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(gcount, &zero[1], sizeof(unsigned), hipMemcpyHostToDevice, x->stream);
        if (e == hipSuccess && !fill) e = cut_run(g, x->stream);
        if (e == hipSuccess) e = cut_compact(g, x->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&cnt[0], x->dcount, sizeof(unsigned), hipMemcpyDeviceToHost, x->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&cnt[1], gcount, sizeof(unsigned), hipMemcpyDeviceToHost, x->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
        return e;
    };
    g.ccap = g.gcap = 0;
    hipError_t e = pass(0);
    int rc = IBTFO_OK;
    if (e == hipSuccess) {
        x->ncut = (int)cnt[0];
        x->ngroups = (int)cnt[1];
        if ((x->ncut && dalloc(x, &x->cuts, x->ncut * sizeof(CutCell)) != hipSuccess) ||
            (x->ngroups && dalloc(x, &x->groups, x->ngroups * sizeof(MergeGroup)) != hipSuccess)) {
            set_error("cut-cell geometry: out of device memory for %d records", x->ncut);
            rc = IBTFO_E_NOMEM;
        } else {
            g.cuts = x->cuts;
            g.ccap = x->ncut;
            g.groups = x->groups;
            g.gcap = x->ngroups;
            e = pass(1);
        }
    }
    dfree(x, dense);
    dfree(x, gcount);
    g.dcut = nullptr;
    if (rc) return rc;
    HIPCHK(e);
    // the merge groups' cells (targets and members) carry MERGED_BIT in the step
    // kernel's type plane: it leaves their dt / mass partials to the merge kernel,
    // which writes one record per block
    dfree(x, x->stype); x->stype = nullptr;
    dfree(x, x->merge_part); x->merge_part = nullptr;
    std::vector<uint8_t> mk(L.plane);
    std::vector<MergeGroup> gr(x->ngroups);
    HIPCHK(scopy(x, mk.data(), x->type, L.plane, hipMemcpyDeviceToHost));
    if (x->ngroups) HIPCHK(scopy(x, gr.data(), x->groups, x->ngroups * sizeof(MergeGroup), hipMemcpyDeviceToHost));
    for (const MergeGroup& m : gr) {
        mk[m.target] |= MERGED_BIT;
        for (int q = 0; q < m.n; q++) mk[m.member[q]] |= MERGED_BIT;
    }
    if (dalloc(x, &x->stype, L.plane) != hipSuccess ||
        (x->ngroups && dalloc(x, &x->merge_part, PART_STRIDE * (size_t)merge_blocks(x->ngroups) * sizeof(double)) !=
                           hipSuccess)) {
        set_error("cut-cell geometry: out of device memory");
        return IBTFO_E_NOMEM;
    }
    HIPCHK(scopy(x, x->stype, mk.data(), L.plane, hipMemcpyHostToDevice));
    return IBTFO_OK;
}

// SimpleBound fills its ghosts in the reference's sequential order (i outer, j
// inner, LS_SimpleBound.cpp:32-36) in place, so a donor that is itself a ghost
// filled earlier is read with its new value. Such dependencies are found here,
// once per geometry: each ghost gets the index of an earlier-filled ghost donor
// (x and y) and a level one above that donor's; the fill then runs level by level.
static int simple_bound_levels(ibtfo_ctx* x) {
    const Layout& L = x->L;
    x->nlevels = 1;
    if (x->cfg.boundary != IBTFO_SIMPLE_BOUND || x->nghost == 0) return IBTFO_OK;
    std::vector<Ghost> gh(x->nghost);
    HIPCHK(scopy(x, gh.data(), x->ghosts, x->nghost * sizeof(Ghost), hipMemcpyDeviceToHost));
    auto key = [&](long long cell) {  // position in the reference's loop order
        const long long r = cell - L.off;
        return (r % L.pitch) * (long long)L.N + r / L.pitch;
    };
    std::vector<int> order(x->nghost);
    std::map<long long, int> at;
    for (int q = 0; q < x->nghost; q++) { order[q] = q; at[gh[q].cell] = q; }
    std::sort(order.begin(), order.end(), [&](int a, int b) { return key(gh[a].cell) < key(gh[b].cell); });
    int top = 0;
    for (int q : order) {
        Ghost& g = gh[q];
        g.pick = 0;
        g.idxe = g.idye = -1;
        auto dep = [&](bool use, long long d, int& slot) {
            if (!use) return;
            auto it = at.find(d);
            if (it == at.end() || key(d) >= key(g.cell)) return;  // not a ghost, or filled later
            slot = it->second;
            g.pick = std::max(g.pick, gh[it->second].pick + 1);
        };
        dep(g.idx != 0, g.d1, g.idxe);
        dep(g.idy != 0, g.d1e, g.idye);
        top = std::max(top, g.pick);
    }
    x->nlevels = top + 1;
    HIPCHK(scopy(x, x->ghosts, gh.data(), x->nghost * sizeof(Ghost), hipMemcpyHostToDevice));
    return IBTFO_OK;
}

// ghost fill phase 1, every dependency level in order
static hipError_t launch_fill_levels(ibtfo_ctx* x, FillArgs f) {
    for (int l = 0; l < std::max(1, x->nlevels); l++) {
        f.level = l;
        hipError_t e = launch_fill(f, x->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int geometry_finish(ibtfo_ctx* x, GeomArgs& g, double* dphi) {
    // the ghost list, cut-cell records and merge groups are rebuilt below (new
    // buffers and counts): a graph captured over the old ones must not replay
    drop_graph(x);
    x->prepared = 0;
    hipError_t ge = geom_run(g, x->stream);
    if (ge != hipSuccess) { dfree(x, dphi); HIPCHK(ge); }
    if (x->cfg.boundary == IBTFO_CUTCELL) {
        int rc = cut_geometry(x, g);
        if (rc) { dfree(x, dphi); return rc; }
    }
    // ghost list (upper bound: all interior cells of type 2; count first)
    dfree(x, x->ghosts); x->ghosts = nullptr;
    dfree(x, x->gvals); x->gvals = nullptr;
    x->nghost = 0;
    if (x->cfg.boundary != IBTFO_REGULAR) {
        unsigned int zero = 0, cnt = 0;
        HIPCHK(hipMemcpyAsync(x->dcount, &zero, sizeof zero, hipMemcpyHostToDevice, x->stream));
        g.count = x->dcount; g.cap = 0; g.ghosts = nullptr;
        HIPCHK(geom_ghosts(g, x->stream));
        HIPCHK(hipMemcpyAsync(&cnt, x->dcount, sizeof cnt, hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
        if (cnt > 0) {
            if (dalloc(x, &x->ghosts, cnt * sizeof(Ghost)) != hipSuccess ||
                dalloc(x, &x->gvals, cnt * 4 * sizeof(double)) != hipSuccess) {
                dfree(x, dphi);
                set_error("geometry: out of device memory for %u ghost records", cnt);
                return IBTFO_E_NOMEM;
            }
            HIPCHK(hipMemcpyAsync(x->dcount, &zero, sizeof zero, hipMemcpyHostToDevice, x->stream));
            g.ghosts = x->ghosts; g.cap = (int)cnt;
            HIPCHK(geom_ghosts(g, x->stream));
        }
        x->nghost = (int)cnt;
    }
    {
        int rc = simple_bound_levels(x);
        if (rc) { dfree(x, dphi); return rc; }
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    dfree(x, dphi);
    x->have_geometry = 1;
    x->prepared = 0;
    return IBTFO_OK;
}

extern "C" int ibtfo_geometry(ibtfo_ctx* x, int kind, const double* prm, int nprm, const double* phi_host) {
    if (!x) return IBTFO_E_ARG;
    const Layout& L = x->L;
    GeomArgs g;
    memset(&g, 0, sizeof g);
    g.L = L;
    g.Li = x->Li; g.Hi = x->Hi; g.dx = x->dxc;
    g.kind = kind;
    g.phi = x->phi; g.nx = x->gnx; g.ny = x->gny; g.type = x->type;
    g.bc = x->cfg.boundary == IBTFO_CUTCELL ? IBTFO_IDW_GFM : x->cfg.boundary;
    g.unit_normals = x->cfg.quirks == 0;
    g.cut = x->cfg.boundary == IBTFO_CUTCELL;
    g.cdx = (x->cfg.hi_x - x->cfg.lo_x) / x->cfg.N_x;
    g.cdy = (x->cfg.hi_y - x->cfg.lo_y) / x->cfg.N_y;
    double* dphi = nullptr;
    auto need = [&](int n) {
        if (nprm < n || !prm) { set_error("level-set hook %d needs %d parameters", kind, n); return false; }
        return true;
    };
    switch (kind) {
        case IBTFO_LS_NONE: break;
        case IBTFO_LS_CHANNEL: {  // Test/InclineSod/Level_Set_function.cpp:35-49, constants on the host
            if (!need(2)) return IBTFO_E_ARG;
            double angle = 3.1415926535 * prm[0] / 180;
            double AxisD = prm[1] / cos(angle);
            double k = tan(angle);
            g.p[0] = k; g.p[1] = pow(1 + k * k, 0.5); g.p[2] = 0.05; g.p[3] = AxisD;
            break;
        }
        case IBTFO_LS_ANNULUS: {  // Test/Vortex/Level_Set_function.cpp:16-20
            if (!need(2)) return IBTFO_E_ARG;
            g.p[0] = prm[0] * 0.40; g.p[1] = prm[1] * 0.40;
            break;
        }
        case IBTFO_LS_ELLIPSE: {
            if (!need(5)) return IBTFO_E_ARG;
            double ang = prm[4] * M_PI / 180.0;
            g.p[0] = prm[0]; g.p[1] = prm[1]; g.p[2] = prm[2]; g.p[3] = prm[3]; g.p[4] = cos(ang); g.p[5] = sin(ang);
            break;
        }
        case IBTFO_LS_SHAPES: {
            if (!prm || nprm < 1 || nprm != 1 + (int)prm[0] * SHAPE_STRIDE) {
                set_error("IBTFO_LS_SHAPES needs 1 + n*%d parameters", SHAPE_STRIDE);
                return IBTFO_E_ARG;
            }
            for (int k = 0; k < (int)prm[0]; k++) {
                const double* sh = prm + 1 + k * SHAPE_STRIDE;
                if (sh[0] != 0.0 && (sh[4] < 3 || sh[4] > 8)) { set_error("polygon needs 3..8 vertices"); return IBTFO_E_ARG; }
            }
            if (nprm > 1) {
                if (dalloc(x, &dphi, (nprm - 1) * sizeof(double)) != hipSuccess) { set_error("geometry: out of device memory"); return IBTFO_E_NOMEM; }
                HIPCHK(scopy(x, dphi, prm + 1, (nprm - 1) * sizeof(double), hipMemcpyHostToDevice));
            }
            g.shapes = dphi;
            g.nshapes = (int)prm[0];
            break;
        }
        case IBTFO_LS_FIELD: {
            if (!phi_host) { set_error("IBTFO_LS_FIELD needs phi"); return IBTFO_E_ARG; }
            if (dalloc(x, &dphi, (size_t)L.M * L.N * sizeof(double)) != hipSuccess) { set_error("geometry: out of device memory"); return IBTFO_E_NOMEM; }
            HIPCHK(scopy(x, dphi, phi_host, (size_t)L.M * L.N * sizeof(double), hipMemcpyHostToDevice));
            g.phi_in = dphi;
            break;
        }
        default: set_error("unknown level-set kind %d", kind); return IBTFO_E_ARG;
    }
    return geometry_finish(x, g, dphi);
}

namespace ibtfo {
int png_read_file(const char* path, std::vector<unsigned char>& rgba, int* width, int* height);
}

// Test/PNGGeo/Level_Set.cpp:11-364 with the image already decoded: phi from the
// mask on the device (mask_run), then normals, classification and the ghost
// list as for any level set.
extern "C" int ibtfo_geometry_mask(ibtfo_ctx* x, const unsigned char* rgba, int width, int height) {
    if (!x || !rgba || width < 3 || height < 3) { set_error("ibtfo_geometry_mask: bad image"); return IBTFO_E_ARG; }
    const Layout& L = x->L;
    const ibtfo_config& c = x->cfg;
    MaskArgs m;
    memset(&m, 0, sizeof m);
    m.W = width; m.H = height; m.M = L.M; m.N = L.N;
    m.Li = x->Li; m.Hi = x->Hi; m.dxc = x->dxc;
    const double Psy_L = c.hi_x - c.lo_x, Psy_H = c.hi_y - c.lo_y;
    m.exo = -c.num_ghost_cell * (Psy_L / c.N_x);  // Level_Set.cpp:82-95
    m.eyo = -c.num_ghost_cell * (Psy_H / c.N_y);
    m.dxp = (Psy_L / c.N_x) * (c.N_x + 2 * c.num_ghost_cell) / width;
    m.dyp = (Psy_H / c.N_y) * (c.N_y + 2 * c.num_ghost_cell) / height;
    const size_t WH = (size_t)width * height, WN = (size_t)width * L.N, NW1 = (size_t)L.N * (width + 1);
    m.nseg = (height + MASK_SEG - 1) / MASK_SEG;
    const size_t SW = (size_t)m.nseg * width;
    // the row envelope runs in LDS up to its limit; wider masks take global scratch
    const bool lds_rows = width < 65536 && mask_lds_bytes(width) <= MASK_LDS_MAX &&
                          hipFuncSetAttribute((const void*)mask_rows_kernel(), hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)mask_lds_bytes(width)) == hipSuccess;
    unsigned char *drgba = nullptr, *dedge = nullptr;
    double *dv = nullptr, *dg1 = nullptr, *dz = nullptr, *dphi = nullptr;
    int *dr1 = nullptr, *dr2 = nullptr, *dvb = nullptr, *dsf = nullptr, *dsl = nullptr, *dbe = nullptr, *daf = nullptr;
    auto release = [&]() {
        for (void* p : {(void*)drgba, (void*)dedge, (void*)dv, (void*)dg1, (void*)dz, (void*)dr1, (void*)dr2,
                        (void*)dvb, (void*)dsf, (void*)dsl, (void*)dbe, (void*)daf})
            dfree(x, p);
    };
    if (dalloc(x, &drgba, WH * 4) != hipSuccess || dalloc(x, &dedge, WH) != hipSuccess ||
        dalloc(x, &dv, WH * sizeof(double)) != hipSuccess || dalloc(x, &dg1, WN * sizeof(double)) != hipSuccess ||
        dalloc(x, &dr1, WN * sizeof(int)) != hipSuccess || dalloc(x, &dr2, WN * sizeof(int)) != hipSuccess ||
        dalloc(x, &dsf, SW * sizeof(int)) != hipSuccess || dalloc(x, &dsl, SW * sizeof(int)) != hipSuccess ||
        dalloc(x, &dbe, SW * sizeof(int)) != hipSuccess || dalloc(x, &daf, SW * sizeof(int)) != hipSuccess ||
        (!lds_rows && (dalloc(x, &dvb, NW1 * sizeof(int)) != hipSuccess ||
                       dalloc(x, &dz, NW1 * sizeof(double)) != hipSuccess)) ||
        dalloc(x, &dphi, (size_t)L.M * L.N * sizeof(double)) != hipSuccess) {
        release();
        dfree(x, dphi);
        set_error("ibtfo_geometry_mask: out of device memory");
        return IBTFO_E_NOMEM;
    }
    m.rgba = drgba; m.v = dv; m.edge = dedge; m.g1 = dg1; m.r1 = dr1; m.r2 = dr2; m.vbuf = dvb; m.zbuf = dz;
    m.sfirst = dsf; m.slast = dsl; m.before = dbe; m.after = daf;
    m.phi = dphi;
    hipError_t e = scopy(x, drgba, rgba, WH * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mask_run(m, x->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
    release();
    if (e != hipSuccess) {
        dfree(x, dphi);
        set_error("ibtfo_geometry_mask: %s", hipGetErrorString(e));
        return IBTFO_E_HIP;
    }
    GeomArgs g;
    memset(&g, 0, sizeof g);
    g.L = L;
    g.Li = x->Li; g.Hi = x->Hi; g.dx = x->dxc;
    g.kind = IBTFO_LS_FIELD;
    g.phi_in = dphi;
    g.phi = x->phi; g.nx = x->gnx; g.ny = x->gny; g.type = x->type;
    g.bc = c.boundary == IBTFO_CUTCELL ? IBTFO_IDW_GFM : c.boundary;
    g.unit_normals = c.quirks == 0;
    g.cut = c.boundary == IBTFO_CUTCELL;
    g.cdx = Psy_L / c.N_x;
    g.cdy = Psy_H / c.N_y;
    return geometry_finish(x, g, dphi);
}

extern "C" int ibtfo_geometry_png(ibtfo_ctx* x, const char* path) {
    if (!x || !path) return IBTFO_E_ARG;
    std::vector<unsigned char> img;
    int w = 0, h = 0;
    int rc = png_read_file(path, img, &w, &h);
    if (rc) return rc;
    return ibtfo_geometry_mask(x, img.data(), w, h);
}

extern "C" int ibtfo_download_geometry(ibtfo_ctx* x, double* XY) {
    if (!x || !XY) return IBTFO_E_ARG;
    const Layout& L = x->L;
    std::vector<double> phi(L.plane), nx(L.plane), ny(L.plane);
    std::vector<uint8_t> t(L.plane);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, phi.data(), x->phi, L.plane * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(scopy(x, nx.data(), x->gnx, L.plane * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(scopy(x, ny.data(), x->gny, L.plane * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(scopy(x, t.data(), x->type, L.plane, hipMemcpyDeviceToHost));
    for (int j = 0; j < L.N; j++)
        for (int i = 0; i < L.M; i++) {
            double* r = XY + ((size_t)L.M * j + i) * 6;
            long long o = L.at(i, j);
            r[0] = x->Li + i * x->dxc;
            r[1] = x->Hi + j * x->dxc;
            r[2] = phi[o]; r[3] = nx[o]; r[4] = ny[o]; r[5] = t[o];
        }
    return IBTFO_OK;
}

extern "C" int ibtfo_download_scheme_index(ibtfo_ctx* x, double* S) {
    if (!x || !S) return IBTFO_E_ARG;
    const Layout& L = x->L;
    const int W = ibtfo_sch_width(x);
    memset(S, 0, sizeof(double) * (size_t)L.M * L.N * W);
    if (x->cfg.boundary == IBTFO_REGULAR || x->nghost == 0) return IBTFO_OK;
    std::vector<Ghost> gh(x->nghost);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, gh.data(), x->ghosts, x->nghost * sizeof(Ghost), hipMemcpyDeviceToHost));
    for (const Ghost& g : gh) {
        long long rel = g.cell - L.off;
        int j = (int)(rel / L.pitch), i = (int)(rel % L.pitch);
        double* r = S + ((size_t)L.M * j + i) * W;
        if (W == 2) {  // SimpleBound: {IDX, IDY}
            r[0] = g.idx; r[1] = g.idy;
            continue;
        }
        r[0] = g.xm; r[1] = g.ym; r[2] = g.idx; r[3] = g.idy;
        if (W >= 8) for (int q = 0; q < 4; q++) r[4 + q] = g.a[q];
        if (W == 16) {
            r[8] = g.xe; r[9] = g.ye; r[10] = g.idxe; r[11] = g.idye;
            for (int q = 0; q < 4; q++) r[12 + q] = g.ae[q];
        }
    }
    return IBTFO_OK;
}

// ------------------------------------------------------------------ state
static void corner_from_aos(ibtfo_ctx* x, const double* U) {
    const Layout& L = x->L;
    x->corner_host.assign(2 * 2 * L.ng * 4, 0.0);
    for (int side = 0; side < 2; side++)
        for (int ri = 0; ri < 2 * L.ng; ri++) {
            int row = ri < L.ng ? ri : L.ny + ri;
            int col = side == 0 ? L.ng : L.nx + L.ng - 1;
            for (int k = 0; k < 4; k++)
                x->corner_host[(side * 2 * L.ng + ri) * 4 + k] = U[((size_t)L.M * row + col) * 4 + k];
        }
}

static int upload_aos_to(ibtfo_ctx* x, const double* U, double* dst) {
    const Layout& L = x->L;
    std::vector<double> planes(4 * L.plane, 0.0);
    for (int j = 0; j < L.N; j++)
        for (int i = 0; i < L.M; i++) {
            const double* q = U + ((size_t)L.M * j + i) * 4;
            long long o = L.at(i, j);
            for (int k = 0; k < 4; k++) planes[o + k * L.plane] = q[k];
        }
    HIPCHK(scopy(x, dst, planes.data(), 4 * L.plane * sizeof(double), hipMemcpyHostToDevice));
    return IBTFO_OK;
}

static int upload_aos(ibtfo_ctx* x, const double* U, int buf) { return upload_aos_to(x, U, x->state[buf]); }

static int read_ctl_head(ibtfo_ctx* x, Ctl& c) {
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, &c, x->ctl, offsetof(Ctl, log_dt), hipMemcpyDeviceToHost));
    return IBTFO_OK;
}
static int write_ctl_head(ibtfo_ctx* x, const Ctl& c) {
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, x->ctl, &c, offsetof(Ctl, log_dt), hipMemcpyHostToDevice));
    return IBTFO_OK;
}

extern "C" int ibtfo_upload_state(ibtfo_ctx* x, const double* U_OLD, const double* U_NEW) {
    if (!x || !U_OLD) return IBTFO_E_ARG;
    HIPCHK(hipStreamSynchronize(x->stream));
    // both buffers: ring cells that no edge policy writes (the PNGGeo hook's
    // corners) keep the uploaded U_OLD in either
    int rc = upload_aos(x, U_OLD, 0);
    if (!rc) rc = upload_aos(x, U_OLD, 1);
    if (rc) return rc;
    if (x->cfg.boundary == IBTFO_SIMPLE_BOUND) {
        if (!x->unew0 && dalloc(x, &x->unew0, 4 * x->L.plane * sizeof(double)) != hipSuccess) {
            set_error("ibtfo_upload_state: out of device memory");
            return IBTFO_E_NOMEM;
        }
        if ((rc = upload_aos_to(x, U_NEW ? U_NEW : U_OLD, x->unew0))) return rc;
    }
    corner_from_aos(x, U_NEW ? U_NEW : U_OLD);
    HIPCHK(scopy(x, x->corner, x->corner_host.data(), x->corner_host.size() * sizeof(double), hipMemcpyHostToDevice));
    Ctl c;
    if ((rc = read_ctl_head(x, c))) return rc;
    c.cur = 0;
    c.ran = 0;
    if ((rc = write_ctl_head(x, c))) return rc;
    x->have_state = 1;
    x->prepared = 0;
    return IBTFO_OK;
}

extern "C" int ibtfo_download_state(ibtfo_ctx* x, double* U) {
    if (!x || !U) return IBTFO_E_ARG;
    const Layout& L = x->L;
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    std::vector<double> planes(4 * L.plane);
    HIPCHK(scopy(x, planes.data(), x->state[c.cur], 4 * L.plane * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < L.N; j++)
        for (int i = 0; i < L.M; i++) {
            double* q = U + ((size_t)L.M * j + i) * 4;
            long long o = L.at(i, j);
            for (int k = 0; k < 4; k++) q[k] = planes[o + k * L.plane];
        }
    return IBTFO_OK;
}

extern "C" int ibtfo_device_state(ibtfo_ctx* x, double** planes4, int* pitch, int* offset) {
    if (!x || !planes4) return IBTFO_E_ARG;
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) planes4[k] = x->state[c.cur] + k * x->L.plane;
    if (pitch) *pitch = x->L.pitch;
    if (offset) *offset = x->L.off;
    return IBTFO_OK;
}

extern "C" int ibtfo_initialize(ibtfo_ctx* x, int kind, const double* prm, int nprm) {
    if (!x) return IBTFO_E_ARG;
    if (!x->have_geometry) { set_error("ibtfo_initialize needs ibtfo_geometry first"); return IBTFO_E_STATE; }
    const Layout& L = x->L;
    if (kind == IBTFO_INIT_UNIFORM && prm && nprm >= 4) {
        // on the device, stream-ordered: the same values as the host initialiser
        // and upload (the conserved state computed here in the same order)
        const double gam = x->cfg.gamma, rho = prm[0], u = prm[1], v = prm[2], p = prm[3];
        const double U[4] = {rho, rho * u, rho * v, p / (gam - 1) + 0.5 * rho * (u * u + v * v)};
        if (x->cfg.boundary == IBTFO_SIMPLE_BOUND && !x->unew0 &&
            dalloc(x, &x->unew0, 4 * L.plane * sizeof(double)) != hipSuccess) {
            set_error("ibtfo_initialize: out of device memory");
            return IBTFO_E_NOMEM;
        }
        HIPCHK(launch_uniform_state(x->state[0], x->state[1],
                                    x->cfg.boundary == IBTFO_SIMPLE_BOUND ? x->unew0 : nullptr, L.plane, U, x->ctl,
                                    x->stream));
        x->corner_host.assign(2 * 2 * L.ng * 4, 0.0);
        for (size_t q = 0; q < x->corner_host.size(); q++) x->corner_host[q] = U[q % 4];
        HIPCHK(hipMemcpyAsync(x->corner, x->corner_host.data(), x->corner_host.size() * sizeof(double),
                              hipMemcpyHostToDevice, x->stream));
        x->have_state = 1;
        x->prepared = 0;
        return IBTFO_OK;
    }
    const size_t cells = (size_t)L.M * L.N;
    std::vector<double> xs(cells), ys(cells), U(cells * 4);
    std::vector<uint8_t> tp(L.plane), t(cells);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, tp.data(), x->type, L.plane, hipMemcpyDeviceToHost));
    for (int j = 0; j < L.N; j++)
        for (int i = 0; i < L.M; i++) {
            size_t q = (size_t)L.M * j + i;
            xs[q] = x->Li + i * x->dxc;
            ys[q] = x->Hi + j * x->dxc;
            t[q] = tp[L.at(i, j)];
        }
    int rc = host_initialize(kind, prm, nprm, L.nx, L.ny, L.ng, x->cfg.gamma, xs.data(), ys.data(), t.data(), U.data());
    if (rc) return rc;
    return ibtfo_upload_state(x, U.data(), nullptr);
}

// -------------------------------------------------------------- time loop
// The deferred finish (one kernel per step) applies when nothing runs between
// the step kernel and its finish: row-parallel strips of a single-step solve on
// the Regular boundary with a ring the step kernel writes (the reference edge
// hook, or the periodic ring), at most DF_MAX_REC blocks, and a single solve
// (its partial records are its own).
static int df_eligible(const ibtfo_ctx* x, int nblocks) {
    const ibtfo_config& c = x->cfg;
    return x->df_part && x->rowpar && c.time == IBTFO_SINGLESTEP && c.boundary == IBTFO_REGULAR &&
           (x->sa.edge_fused == 1 || x->sa.edge_fused == 2) && nblocks <= DF_MAX_REC;
}

// An RK2 step needs no kernel between its stages when the step kernel writes the
// ring itself (the reference ring, whose corners on the Regular boundary come
// from the fixed corner table, or the periodic ring) and there is no ghost fill
// (Regular boundary): stage 2 then reads U1 from the buffer stage 1 wrote, with
// no finish making it current.
static bool rk2_in_place(const ibtfo_ctx* x) {
    return x->cfg.boundary == IBTFO_REGULAR && (x->sa.edge_fused == 1 || x->sa.edge_fused == 2);
}

void ibtfo::build_args(ibtfo_ctx* x) {
    const ibtfo_config& c = x->cfg;
    const Layout& L = x->L;
    StepArgs& a = x->sa;
    memset(&a, 0, sizeof a);
    a.buf[0] = x->state[0]; a.buf[1] = x->state[1];
    a.L = L;
    a.type = x->type;
    a.ctl = x->ctl;
    a.mass_part = x->mass_part;
    a.corner = x->corner;
    a.gam = c.gamma; a.gm1 = c.gamma - 1; a.gp1 = c.gamma + 1; a.c1 = (c.gamma - 1) / 2 / c.gamma;
    a.dx = (c.hi_x - c.lo_x) / c.N_x;   // MUSCL_Scheme.cpp:64-65
    a.dy = (c.hi_y - c.lo_y) / c.N_y;
    a.r_gam = 1.0 / a.gam; a.r_gm1 = 1.0 / a.gm1; a.r_dx = 1.0 / a.dx; a.r_dy = 1.0 / a.dy;
    a.cfl = c.CFL_number;
    a.t_end = c.Psy_time;
    a.max_iter = c.max_iter;
    a.quirks = c.quirks;
    a.edge_fused = (c.edge == IBTFO_EDGE_REFERENCE && c.scheme != IBTFO_LAX) ? 1 : 0;
    // a periodic ring with no body is a copy of the new interior: the step kernel
    // writes it beside its stores, and no edge kernel runs
    if (c.edge == IBTFO_EDGE_PERIODIC && c.boundary == IBTFO_REGULAR && c.scheme != IBTFO_LAX && L.nx >= L.ng &&
        L.ny >= L.ng)
        a.edge_fused = 2;
    // the reference's U_NEW ring is refreshed by Boundary's copy back (GFM family),
    // else it keeps the uploaded values (Regular, SimpleBound)
    a.corner_from_input = c.boundary != IBTFO_REGULAR && c.boundary != IBTFO_SIMPLE_BOUND;
    a.unew_nf = c.boundary == IBTFO_SIMPLE_BOUND ? x->unew0 : nullptr;
    step_grid(L, x->rows_override, &a.nbx, &a.nby, &a.rows);
    a.rowpar = x->rowpar;
    a.cut = x->cfg.boundary == IBTFO_CUTCELL ? x->cuts : nullptr;
    a.cut_index = x->cut_index;
    if (a.cut) a.type = x->stype;
    // deferred finish: a step is one kernel when nothing runs between the step
    // kernel and its finish (no ghost fill, edge or merge kernel)
    a.df_part = x->df_part;
    a.df_ok = df_eligible(x, a.nbx * a.nby);
    a.df = 0;
    a.df_k = 0;

    FillArgs& f = x->fa;
    memset(&f, 0, sizeof f);
    f.buf[0] = x->state[0]; f.buf[1] = x->state[1];
    f.L = L; f.ctl = x->ctl; f.ghosts = x->ghosts; f.nghost = x->nghost; f.out = x->gvals;
    f.bc = c.boundary == IBTFO_CUTCELL ? IBTFO_IDW_GFM : c.boundary; f.gam = c.gamma; f.gm1 = c.gamma - 1;
    f.dxg = (x->Li + 1 * x->dxc) - (x->Li + 0 * x->dxc);
    f.dyg = (x->Hi + 2 * x->dxc) - (x->Hi + 1 * x->dxc);
    f.use_cur = 1; f.require_ran = 1; f.ring_prev = 1; f.quirks = c.quirks;

    EdgeArgs& e = x->ea;
    memset(&e, 0, sizeof e);
    e.buf[0] = x->state[0]; e.buf[1] = x->state[1];
    e.L = L; e.ctl = x->ctl; e.kind = c.edge; e.gam = c.gamma; e.require_ran = 1;
    e.unew_nf = a.unew_nf;
    e.type = x->type;
    double r = c.edge_state[0], u = c.edge_state[1], v = c.edge_state[2], p = c.edge_state[3];
    e.inflow[0] = r; e.inflow[1] = r * u; e.inflow[2] = r * v;
    e.inflow[3] = p / (c.gamma - 1) + 0.5 * r * (u * u + v * v);

    FinishArgs& fn = x->fin;
    memset(&fn, 0, sizeof fn);
    fn.ctl = x->ctl; fn.part = x->mass_part; fn.nblocks = a.nbx * a.nby;
    fn.cfl = c.CFL_number; fn.dx = a.dx; fn.t_end = c.Psy_time; fn.max_iter = c.max_iter;
    fn.advance = 1; fn.reset = 0;
    fn.keep_cur = c.time == IBTFO_RK2 && rk2_in_place(x);
    fn.fold = x->fold;
    if (c.boundary == IBTFO_CUTCELL && x->ngroups) {  // the merge kernel's records follow the step kernel's
        fn.part2 = x->merge_part;
        fn.nblocks2 = merge_blocks(x->ngroups);
    }

    ReduceArgs& q = x->ra;
    memset(&q, 0, sizeof q);
    q.L = L; q.type = x->type; q.ctl = x->ctl; q.mass_part = x->mass_part;
    q.gam = c.gamma; q.gm1 = c.gamma - 1; q.dx = a.dx; q.dy = a.dy; q.cfl = c.CFL_number;
    q.t_end = c.Psy_time; q.max_iter = c.max_iter;
    q.nblocks = std::min(x->mass_cap, std::max(1, (L.nx * L.ny + 255) / 256));
    q.cut = a.cut;
    q.cut_index = x->cut_index;

    MergeArgs& mg = x->ma;
    memset(&mg, 0, sizeof mg);
    mg.buf[0] = x->state[0]; mg.buf[1] = x->state[1];
    mg.plane = L.plane;
    mg.ctl = x->ctl; mg.groups = x->groups; mg.ngroups = x->ngroups;
    mg.gm1 = c.gamma - 1; mg.r_gm1 = 1.0 / (c.gamma - 1); mg.quirks = c.quirks;
    mg.dx = a.dx; mg.dy = a.dy;
    mg.part = x->merge_part;
}

// Boundary and Global_Boundary after an update (the state just made current):
// ghost fill phase 1, the edge ring, ghost fill phase 2.
static int enqueue_boundaries(ibtfo_ctx* x) {
    const ibtfo_config& c = x->cfg;
    if (c.boundary != IBTFO_REGULAR && c.edge != IBTFO_EDGE_REFERENCE) {
        // fill (first level) and ring in one launch; further SimpleBound levels after
        FillArgs f = x->fa;
        f.level = 0;
        HIPCHK(launch_fill_edge(f, x->ea, x->stream));
        for (int l = 1; l < x->nlevels; l++) {
            f.level = l;
            HIPCHK(launch_fill(f, x->stream));
        }
        HIPCHK(launch_scatter(x->fa, x->stream));
        return IBTFO_OK;
    }
    if (c.boundary != IBTFO_REGULAR) HIPCHK(launch_fill_levels(x, x->fa));
    if (c.edge == IBTFO_EDGE_REFERENCE && c.scheme == IBTFO_LAX)
        HIPCHK(launch_edge_ref(x->ea, x->corner, x->sa.corner_from_input, x->stream));
    if (c.edge != IBTFO_EDGE_REFERENCE && x->sa.edge_fused != 2) HIPCHK(launch_edge(x->ea, x->stream));
    if (c.boundary != IBTFO_REGULAR) HIPCHK(launch_scatter(x->fa, x->stream));
    return IBTFO_OK;
}

// One step of the reference loop. `ev` (optional) receives events around each
// step-kernel launch: ev[0], ev[1] (single step or RK2 stage 1), ev[2], ev[3]
// (RK2 stage 2). k > 0: the k-th step of a deferred-finish batch (one kernel;
// the next kernel, or the batch's flush, finishes it).
static int enqueue_step(ibtfo_ctx* x, hipEvent_t* ev = nullptr, int k = 0) {
    const ibtfo_config& c = x->cfg;
    int rc;
    if (c.time == IBTFO_RK2) {
        // Heun: U1 = U^n + dt L(U^n) into the other buffer, made current, with its
        // boundaries; then (U^n + U1 + dt L(U1)) / 2 written over U^n. With nothing
        // to run between the stages, stage 2 reads U1 from the other buffer (stage
        // 3) instead of a finish making it current, and the step's finish keeps cur.
        const bool in_place = rk2_in_place(x);
        StepArgs s1 = x->sa;
        s1.stage = 1;
        if (ev) HIPCHK(hipEventRecord(ev[0], x->stream));
        HIPCHK(launch_step(c.scheme, c.riemann, c.limiter, c.precision, s1, x->stream));
        if (ev) HIPCHK(hipEventRecord(ev[1], x->stream));
        if (!in_place) {
            FinishArgs flip = x->fin;
            flip.advance = 2;
            HIPCHK(launch_finish(flip, x->stream));
            if ((rc = enqueue_boundaries(x))) return rc;
        }
        StepArgs s2 = x->sa;
        s2.stage = in_place ? 3 : 2;
        if (ev) HIPCHK(hipEventRecord(ev[2], x->stream));
        HIPCHK(launch_step(c.scheme, c.riemann, c.limiter, c.precision, s2, x->stream));
        if (ev) HIPCHK(hipEventRecord(ev[3], x->stream));
    } else {
        StepArgs s = x->sa;
        if (k > 0) {
            s.df = k == 1 ? 1 : 2;
            s.df_k = k & 1;
        }
        if (ev) HIPCHK(hipEventRecord(ev[0], x->stream));
        HIPCHK(launch_step(c.scheme, c.riemann, c.limiter, c.precision, s, x->stream));
        if (ev) HIPCHK(hipEventRecord(ev[1], x->stream));
        if (k > 0) return IBTFO_OK;
    }
    if (c.boundary == IBTFO_CUTCELL) {
        // small cells merged (with the merged cells' dt / mass partials), then
        // ComputeDt / mass over both kernels' partials
        HIPCHK(launch_merge(x->ma, x->stream));
    }
    HIPCHK(launch_finish(x->fin, x->stream));
    return enqueue_boundaries(x);
}

// n steps, queued as one batch: with the deferred finish, n step kernels and
// the closing flush, which leaves the control block as the finish kernel would.
// `ev` (optional): `per` events per step, as enqueue_step takes them.
static int enqueue_steps(ibtfo_ctx* x, int n, hipEvent_t* ev = nullptr, int per = 0) {
    const bool df = x->sa.df_ok != 0;
    for (int s = 0; s < n; s++) {
        int rc = enqueue_step(x, ev ? ev + (size_t)per * s : nullptr, df ? s + 1 : 0);
        if (rc) return rc;
    }
    if (df && n > 0) HIPCHK(launch_df_flush(x->sa, n & 1, x->stream));
    return IBTFO_OK;
}

static int run_reduce(ibtfo_ctx* x, int reset) {
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    x->ra.U = x->state[c.cur];
    x->ra.reset = reset;
    HIPCHK(launch_reduce(x->ra, x->stream));
    return IBTFO_OK;
}

extern "C" int ibtfo_boundary(ibtfo_ctx* x) {
    if (!x) return IBTFO_E_ARG;
    if (!x->have_geometry || !x->have_state) { set_error("boundary needs geometry and state"); return IBTFO_E_STATE; }
    build_args(x);
    if (x->cfg.boundary == IBTFO_REGULAR) return IBTFO_OK;  // U_OLD = U_NEW on the interior: already equal
    FillArgs f = x->fa;
    f.require_ran = 0;
    f.ring_prev = 0;  // U_NEW = U_OLD here, ring included
    HIPCHK(launch_fill_levels(x, f));
    HIPCHK(launch_scatter(f, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return IBTFO_OK;
}

extern "C" int ibtfo_prepare(ibtfo_ctx* x, int apply_boundary) {
    if (!x) return IBTFO_E_ARG;
    if (!x->have_geometry || !x->have_state) { set_error("prepare needs geometry and state"); return IBTFO_E_STATE; }
    build_args(x);
    int rc;
    if (apply_boundary && (rc = ibtfo_boundary(x))) return rc;
    if ((rc = run_reduce(x, 1))) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    x->prepared = 1;
    return IBTFO_OK;
}

extern "C" int ibtfo_step(ibtfo_ctx* x, int nsteps) {
    if (!x || nsteps < 0) return IBTFO_E_ARG;
    if (!x->prepared) { set_error("ibtfo_step before ibtfo_prepare"); return IBTFO_E_STATE; }
    return enqueue_steps(x, nsteps);
}

extern "C" int ibtfo_step_dt(ibtfo_ctx* x, double dt) {
    if (!x) return IBTFO_E_ARG;
    if (!(dt > 0.0) || !std::isfinite(dt)) { set_error("ibtfo_step_dt: dt = %g", dt); return IBTFO_E_ARG; }
    if (!x->prepared) { set_error("ibtfo_step_dt before ibtfo_prepare"); return IBTFO_E_STATE; }
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    c.dt = dt;   // the caller's dt for this step
    c.done = 0;  // the caller keeps the clock and decides when to stop
    if ((rc = write_ctl_head(x, c))) return rc;
    return enqueue_steps(x, 1);
}

extern "C" int ibtfo_step_graph(ibtfo_ctx* x, int nsteps, int chunk) {
    if (!x || nsteps < 0 || chunk <= 0) return IBTFO_E_ARG;
    if (!x->prepared) { set_error("ibtfo_step_graph before ibtfo_prepare"); return IBTFO_E_STATE; }
    if (!x->graph || x->graph_chunk != chunk) {
        if (x->graph) { (void)hipGraphExecDestroy(x->graph); x->graph = nullptr; }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
        int rc = IBTFO_OK;
        rc = enqueue_steps(x, chunk);
        hipError_t e = hipStreamEndCapture(x->stream, &g);
        if (e == hipSuccess && rc == IBTFO_OK && g) e = hipGraphInstantiate(&x->graph, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        HIPCHK(e);
        x->graph_chunk = chunk;
    }
    int full = nsteps / chunk, rest = nsteps % chunk;
    for (int r = 0; r < full; r++) HIPCHK(hipGraphLaunch(x->graph, x->stream));
    return ibtfo_step(x, rest);
}

extern "C" int ibtfo_status(ibtfo_ctx* x, int* iter, double* now_t, double* dt) {
    if (!x) return IBTFO_E_ARG;
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    if (iter) *iter = c.iter;
    if (now_t) *now_t = c.now_t;
    if (dt) *dt = c.dt;
    return c.status ? IBTFO_E_NONFINITE : IBTFO_OK;
}

extern "C" int ibtfo_log(ibtfo_ctx* x, int count, double* dt, double* t, double* mass) {
    if (!x || count < 0 || count > LOGCAP - 1) return IBTFO_E_ARG;
    Ctl c;
    int rc = read_ctl_head(x, c);
    if (rc) return rc;
    if (count > c.iter) { set_error("only %d steps logged", c.iter); return IBTFO_E_ARG; }
    std::vector<double> ld(LOGCAP), lt(LOGCAP), lm(LOGCAP);
    HIPCHK(scopy(x, ld.data(), x->ctl->log_dt, sizeof(double) * LOGCAP, hipMemcpyDeviceToHost));
    HIPCHK(scopy(x, lt.data(), x->ctl->log_t, sizeof(double) * LOGCAP, hipMemcpyDeviceToHost));
    HIPCHK(scopy(x, lm.data(), x->ctl->log_mass, sizeof(double) * LOGCAP, hipMemcpyDeviceToHost));
    for (int q = 0; q < count; q++) {
        int it = c.iter - count + q;
        int s = it % LOGCAP;
        if (dt) dt[q] = ld[s];
        if (t) t[q] = lt[s];
        if (mass) mass[q] = lm[s];
    }
    return IBTFO_OK;
}

extern "C" int ibtfo_compute_dt(ibtfo_ctx* x, double* dt) {
    if (!x || !dt) return IBTFO_E_ARG;
    if (!x->have_geometry || !x->have_state) return IBTFO_E_STATE;
    build_args(x);
    int rc = run_reduce(x, 0);
    if (rc) return rc;
    Ctl c;
    if ((rc = read_ctl_head(x, c))) return rc;
    *dt = c.dt;
    return IBTFO_OK;
}

extern "C" int ibtfo_mass(ibtfo_ctx* x, double* mass) {
    if (!x || !mass) return IBTFO_E_ARG;
    if (!x->have_geometry || !x->have_state) return IBTFO_E_STATE;
    build_args(x);
    int rc = run_reduce(x, 0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(scopy(x, mass, &x->ctl->mass_now, sizeof(double), hipMemcpyDeviceToHost));
    return IBTFO_OK;
}

extern "C" int ibtfo_time_steps(ibtfo_ctx* x, int nsteps, double* total_ms, double* step_kernel_ms) {
    if (!x || nsteps <= 0) return IBTFO_E_ARG;
    if (!x->prepared) return IBTFO_E_STATE;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipEventRecord(e0, x->stream));
    int rc = x->graph ? ibtfo_step_graph(x, nsteps, x->graph_chunk) : ibtfo_step(x, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, x->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (total_ms) *total_ms = ms;
    if (step_kernel_ms) {
        // per-launch events around the step kernel(s) of each step
        const int per = x->cfg.time == IBTFO_RK2 ? 4 : 2;
        std::vector<hipEvent_t> ev((size_t)per * nsteps, nullptr);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        rc = enqueue_steps(x, nsteps, ev.data(), per);
        double sum = 0;
        if (rc == IBTFO_OK && hipEventSynchronize(ev.back()) != hipSuccess) rc = IBTFO_E_HIP;
        for (int s = 0; s < nsteps && rc == IBTFO_OK; s++)
            for (int q = 0; q < per; q += 2) {
                float m = 0;
                if (hipEventElapsedTime(&m, ev[(size_t)per * s + q], ev[(size_t)per * s + q + 1]) != hipSuccess) {
                    rc = IBTFO_E_HIP;
                    break;
                }
                sum += m;
            }
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (rc) { set_error("ibtfo_time_steps: per-launch timing failed"); return rc; }
        *step_kernel_ms = sum / nsteps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return IBTFO_OK;
}

extern "C" int ibtfo_set_tile(ibtfo_ctx* x, int rows, int row_parallel) {
    if (!x || rows < 0 || (row_parallel != 0 && row_parallel != 1)) {
        set_error("ibtfo_set_tile: bad argument");
        return IBTFO_E_ARG;
    }
    if (row_parallel && (rows < 1 || rows > ROWPAR_MAX_ROWS || x->cfg.scheme == IBTFO_LAX)) {
        set_error("ibtfo_set_tile: the row-parallel form takes 1..%d rows and no Lax scheme (rows %d)",
                  ROWPAR_MAX_ROWS, rows);
        return IBTFO_E_ARG;
    }
    int nbx, nby, r;
    step_grid(x->L, rows, &nbx, &nby, &r);
    if (nbx * nby > x->mass_cap) {
        double* p = nullptr;
        if (dalloc(x, &p, PART_STRIDE * (size_t)nbx * nby * sizeof(double)) != hipSuccess) {
            set_error("ibtfo_set_tile: allocation of %d block partials failed", nbx * nby);
            return IBTFO_E_NOMEM;
        }
        HIPCHK(hipStreamSynchronize(x->stream));
        dfree(x, x->mass_part);
        x->mass_part = p;
        x->mass_cap = nbx * nby;
    }
    x->rows_override = rows;
    x->rowpar = row_parallel;
    drop_graph(x);
    build_args(x);
    return IBTFO_OK;
}

extern "C" int ibtfo_set_rows(ibtfo_ctx* x, int rows) { return ibtfo_set_tile(x, rows, 0); }

extern "C" int ibtfo_tile(const ibtfo_ctx* x, int* rows, int* nbx, int* nby) {
    if (!x) return IBTFO_E_ARG;
    int a, b, r;
    step_grid(x->L, x->rows_override, &a, &b, &r);
    if (rows) *rows = r;
    if (nbx) *nbx = a;
    if (nby) *nby = b;
    return IBTFO_OK;
}

extern "C" int ibtfo_tile_ex(const ibtfo_ctx* x, int* rows, int* nbx, int* nby, int* row_parallel) {
    if (!x) return IBTFO_E_ARG;
    if (row_parallel) *row_parallel = x->rowpar;
    return ibtfo_tile(x, rows, nbx, nby);
}

extern "C" int ibtfo_one_kernel_steps(const ibtfo_ctx* x) {
    if (!x) return -1;
    int nbx, nby, r;
    step_grid(x->L, x->rows_override, &nbx, &nby, &r);
    return df_eligible(x, nbx * nby) ? 1 : 0;
}

extern "C" int ibtfo_autotune(ibtfo_ctx* x, int probe_steps, int* best_rows, double* best_ms) {
    if (!x || probe_steps <= 0) { set_error("ibtfo_autotune: bad argument"); return IBTFO_E_ARG; }
    if (!x->prepared) { set_error("ibtfo_autotune before ibtfo_prepare"); return IBTFO_E_STATE; }
    Ctl head;
    int rc = read_ctl_head(x, head);
    if (rc) return rc;
    if (head.done) { set_error("ibtfo_autotune: the run has already finished"); return IBTFO_E_STATE; }
    const Layout& L = x->L;
    const size_t sb = 4 * (size_t)L.plane * sizeof(double);
    char* snap = nullptr;
    if (hipMalloc(&snap, 2 * sb + sizeof(Ctl)) != hipSuccess) {
        set_error("ibtfo_autotune: allocation of the state snapshot failed");
        return IBTFO_E_NOMEM;
    }
    auto copy_all = [&](bool save) -> int {
        char* st0 = snap;
        char* st1 = snap + sb;
        char* ct = snap + 2 * sb;
        HIPCHK(hipMemcpyAsync(save ? st0 : (char*)x->state[0], save ? (char*)x->state[0] : st0, sb,
                              hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(save ? st1 : (char*)x->state[1], save ? (char*)x->state[1] : st1, sb,
                              hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(save ? ct : (char*)x->ctl, save ? (char*)x->ctl : ct, sizeof(Ctl),
                              hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
        return IBTFO_OK;
    };
    const int keep = x->rows_override, keep_par = x->rowpar;
    // short strips for small grids (latency: a wave's rows run one after the
    // other), tall ones for large grids (fewer halo rows re-read per row updated);
    // then the row-parallel form (each row of a strip on its own wave)
    struct Cand { int rows, par; };
    std::vector<Cand> cand;
    for (int r : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256}) cand.push_back({r, 0});
    if (x->cfg.scheme != IBTFO_LAX)
        for (int r = 1; r <= ROWPAR_MAX_ROWS; r++) cand.push_back({r, 1});
    int best = 0, best_par = 0;
    double bms = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = copy_all(true);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        set_error("ibtfo_autotune: event creation failed");
        rc = IBTFO_E_HIP;
    }
    for (const Cand& cd : cand) {
        if (rc) break;
        const int r = cd.rows;
        if (r > L.ny) continue;
        if ((rc = ibtfo_set_tile(x, r, cd.par))) break;
        if ((rc = ibtfo_step(x, 2))) break;
        if (hipEventRecord(e0, x->stream) != hipSuccess) { rc = IBTFO_E_HIP; break; }
        if ((rc = ibtfo_step(x, probe_steps))) break;
        float ms = 0;
        if (hipEventRecord(e1, x->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            set_error("ibtfo_autotune: timing failed");
            rc = IBTFO_E_HIP;
            break;
        }
        Ctl h;
        if ((rc = read_ctl_head(x, h))) break;
        if (h.done) { set_error("ibtfo_autotune: the run ends within the probe"); rc = IBTFO_E_STATE; break; }
        if (ms / probe_steps < bms) { bms = ms / probe_steps; best = r; best_par = cd.par; }
        if ((rc = copy_all(false))) break;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        (void)copy_all(false);
        (void)hipFree(snap);
        (void)ibtfo_set_tile(x, keep, keep_par);
        return rc;
    }
    (void)hipFree(snap);
    if ((rc = ibtfo_set_tile(x, best, best_par))) return rc;
    if (best_rows) *best_rows = best;
    if (best_ms) *best_ms = bms;
    return IBTFO_OK;
}
