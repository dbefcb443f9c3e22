// Batch engine: many independent solves served by one device-resident loop
// (include/ibtfo.h, "Batch engine"). Solves of different sizes share batched
// launches of the step, finish, ghost-fill and edge kernels: each launch's grid
// covers the strips (ghosts, ring cells) of every resident solve through a
// block map, and the whole chunk of steps is one captured HIP graph whose
// arguments never change (tables and maps live at fixed device addresses and
// the grids are sized for the batch's capacity), so solves come and go between
// replays without re-capturing. Device memory of admitted solves comes from an
// arena reserved once; admission (geometry, initial state) runs on a second
// stream while the previous chunk executes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "ibtfo_ctx.h"

using namespace ibtfo;

// ------------------------------------------------------------------ arena
namespace ibtfo {
static const size_t ARENA_ALIGN = 256;
// admissions per progress call while no chunk is running (the next call starts one)
static const int ADMIT_BURST = 8;

int Arena::init(size_t bytes) {
    cap_ = bytes / ARENA_ALIGN * ARENA_ALIGN;
    if (cap_ == 0) { set_error("arena of zero bytes"); return IBTFO_E_ARG; }
    if (hipMalloc((void**)&base_, cap_) != hipSuccess) {
        base_ = nullptr;
        set_error("cannot reserve %zu bytes of device memory for the batch", cap_);
        return IBTFO_E_NOMEM;
    }
    free_.clear();
    live_.clear();
    free_[0] = cap_;
    used_ = 0;
    return IBTFO_OK;
}
void Arena::release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = used_ = 0;
    free_.clear();
    live_.clear();
}
void* Arena::alloc(size_t bytes) {
    size_t n = std::max<size_t>(ARENA_ALIGN, (bytes + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > n) free_[off + n] = sz - n;
        live_[off] = n;
        used_ += n;
        return base_ + off;
    }
    return nullptr;
}
void Arena::free(void* p) {
    if (!p) return;
    size_t off = (size_t)((char*)p - base_);
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t sz = lv->second;
    live_.erase(lv);
    used_ -= sz;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + sz == nx->first) {  // merge with the next free range
        sz += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {  // merge with the previous one
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += sz;
            return;
        }
    }
    free_[off] = sz;
}
size_t Arena::largest_free() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
}
}  // namespace ibtfo

// ------------------------------------------------------------------ batch
namespace {
struct Req {
    long long ticket;
    ibtfo_config cfg;
    int ls_kind, init_kind, keep;
    std::vector<double> ls_prm, phi, init_prm;
    double arrival;
};
struct Done {
    ibtfo_result r;
    std::vector<double> U;
};
struct Slot {
    ibtfo_ctx* ctx = nullptr;
    Req req;
    double t_admit = 0;
    long long cells = 0;
};
template <class T> static hipError_t dev_alloc(T** p, size_t n) { return hipMalloc((void**)p, n * sizeof(T)); }
template <class T> static hipError_t host_alloc(T** p, size_t n) { return hipHostMalloc((void**)p, n * sizeof(T), 0); }
}  // namespace

struct ibtfo_batch {
    ibtfo_batch_config bc;
    Arena arena;
    hipStream_t run = nullptr, adm = nullptr;
    std::chrono::steady_clock::time_point t0;
    std::vector<Slot> slots;
    std::vector<std::pair<ibtfo_ctx*, Req>> pending;  // admitted, not yet in the tables
    std::vector<double> pending_t;
    std::deque<Req> queue;
    std::map<long long, Done> done;
    std::deque<long long> ready;
    long long next_ticket = 1;
    // device tables at fixed addresses (the captured graph refers to them) and pinned mirrors
    StepArgs *d_step = nullptr, *h_step = nullptr;
    FinishArgs *d_fin = nullptr, *h_fin = nullptr;
    FillArgs *d_fill = nullptr, *h_fill = nullptr;
    EdgeArgs *d_edge = nullptr, *h_edge = nullptr;
    int2 *d_smap = nullptr, *h_smap = nullptr, *d_gmap = nullptr, *h_gmap = nullptr, *d_rmap = nullptr,
         *h_rmap = nullptr;
    SlotStatus *d_stat = nullptr, *h_stat = nullptr;
    // persistent mode: item counts the kernel reads, its barrier counter and error flag
    int *d_counts = nullptr, *h_counts = nullptr, *d_err = nullptr;
    unsigned int* d_bar = nullptr;
    int persist_grid = 0;
    int used_step = 0, used_ghost = 0, used_ring = 0;  // blocks of the resident solves
    hipGraphExec_t graph = nullptr;
    bool dirty = true;
    long long chunks = 0, steps = 0, cell_updates = 0;
    int finished = 0;
    double busy = 0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return IBTFO_E_HIP;                                                    \
        }                                                                          \
    } while (0)

static int step_blocks(const ibtfo_ctx* x) { return x->sa.nbx * x->sa.nby; }
static int ghost_blocks(const ibtfo_ctx* x) { return (x->nghost + FILL_THREADS - 1) / FILL_THREADS; }
static int ring_blocks(const ibtfo_ctx* x) {
    if (x->cfg.edge == IBTFO_EDGE_REFERENCE) return 0;  // written by the step kernel
    const int nring = x->L.M * x->L.N - x->L.nx * x->L.ny;
    return (nring + EDGE_THREADS - 1) / EDGE_THREADS;
}

// The same counts from a request's configuration, before its context exists
// (the ghost count needs the geometry and is checked at admission).
static int cfg_step_blocks(const ibtfo_config& c, int rows) {
    const int nbx = (c.N_x + STRIP_W - 1) / STRIP_W;
    const int r = std::max(1, std::min(rows, c.N_y));
    return nbx * ((c.N_y + r - 1) / r);
}
static int cfg_ring_blocks(const ibtfo_config& c) {
    if (c.edge == IBTFO_EDGE_REFERENCE) return 0;
    const long long M = c.N_x + 2 * c.num_ghost_cell, N = c.N_y + 2 * c.num_ghost_cell;
    return (int)((M * N - (long long)c.N_x * c.N_y + EDGE_THREADS - 1) / EDGE_THREADS);
}

extern "C" int ibtfo_batch_config_default(ibtfo_batch_config* bc) {
    if (!bc) return IBTFO_E_ARG;
    memset(bc, 0, sizeof *bc);
    bc->scheme = IBTFO_MUSCL;
    bc->riemann = IBTFO_HLL;
    bc->limiter = IBTFO_MINMOD;
    bc->boundary = IBTFO_REGULAR;
    bc->time = IBTFO_SINGLESTEP;
    bc->precision = IBTFO_FP64;
    bc->quirks = 1;
    bc->hbm_budget = 8ll << 30;
    bc->max_slots = 256;
    bc->max_step_blocks = 65536;
    bc->max_ghost_blocks = 16384;
    bc->max_ring_blocks = 16384;
    bc->rows = 16;
    bc->chunk = 50;
    bc->persistent = 0;
    return IBTFO_OK;
}

static void batch_free(ibtfo_batch* b) {
    if (!b) return;
    if (b->run) (void)hipStreamSynchronize(b->run);
    if (b->adm) (void)hipStreamSynchronize(b->adm);
    for (auto& s : b->slots)
        if (s.ctx) free_ctx(s.ctx);
    for (auto& p : b->pending) free_ctx(p.first);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (void* p : {(void*)b->d_step, (void*)b->d_fin, (void*)b->d_fill, (void*)b->d_edge, (void*)b->d_smap,
                    (void*)b->d_gmap, (void*)b->d_rmap, (void*)b->d_stat, (void*)b->d_counts, (void*)b->d_bar,
                    (void*)b->d_err})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)b->h_step, (void*)b->h_fin, (void*)b->h_fill, (void*)b->h_edge, (void*)b->h_smap,
                    (void*)b->h_gmap, (void*)b->h_rmap, (void*)b->h_stat, (void*)b->h_counts})
        if (p) (void)hipHostFree(p);
    b->arena.release();
    if (b->run) (void)hipStreamDestroy(b->run);
    if (b->adm) (void)hipStreamDestroy(b->adm);
    delete b;
}

static int ensure_persistent(ibtfo_batch* b);

extern "C" int ibtfo_batch_create(const ibtfo_batch_config* bc, ibtfo_batch** out) {
    if (!bc || !out) { set_error("null argument"); return IBTFO_E_ARG; }
    *out = nullptr;
    if (bc->max_slots <= 0 || bc->max_step_blocks <= 0 || bc->max_ghost_blocks < 0 || bc->max_ring_blocks < 0 ||
        bc->chunk <= 0 || bc->rows < 0 || bc->hbm_budget <= 0) {
        set_error("ibtfo_batch_create: bad capacities");
        return IBTFO_E_ARG;
    }
    if (bc->scheme == IBTFO_LAX) { set_error("the batch engine runs the reconstructing schemes only"); return IBTFO_E_ARG; }
    if (bc->boundary == IBTFO_CUTCELL) { set_error("the batch engine has no cut-cell step sequence"); return IBTFO_E_ARG; }
    if (bc->boundary == IBTFO_SIMPLE_BOUND) {
        set_error("the batch engine runs one ghost-fill pass per step; SimpleBound may need several");
        return IBTFO_E_ARG;
    }
    ibtfo_batch* b = new ibtfo_batch();
    b->bc = *bc;
    if (b->bc.rows == 0) b->bc.rows = 16;
    b->t0 = std::chrono::steady_clock::now();
    const int S = bc->max_slots;
    int rc = b->arena.init((size_t)bc->hbm_budget);
    if (rc) { batch_free(b); return rc; }
    if (hipStreamCreateWithFlags(&b->run, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->adm, hipStreamNonBlocking) != hipSuccess || dev_alloc(&b->d_step, S) ||
        dev_alloc(&b->d_fin, S) || dev_alloc(&b->d_fill, S) || dev_alloc(&b->d_edge, S) ||
        dev_alloc(&b->d_stat, S) || dev_alloc(&b->d_smap, bc->max_step_blocks) ||
        dev_alloc(&b->d_gmap, std::max(1, bc->max_ghost_blocks)) ||
        dev_alloc(&b->d_rmap, std::max(1, bc->max_ring_blocks)) || host_alloc(&b->h_step, S) ||
        host_alloc(&b->h_fin, S) || host_alloc(&b->h_fill, S) || host_alloc(&b->h_edge, S) ||
        host_alloc(&b->h_stat, S) || host_alloc(&b->h_smap, bc->max_step_blocks) ||
        host_alloc(&b->h_gmap, std::max(1, bc->max_ghost_blocks)) ||
        host_alloc(&b->h_rmap, std::max(1, bc->max_ring_blocks))) {
        set_error("ibtfo_batch_create: allocation of streams or tables failed");
        batch_free(b);
        return IBTFO_E_NOMEM;
    }
    b->slots.resize(S);
    if (bc->persistent && (rc = ensure_persistent(b))) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_destroy(ibtfo_batch* b) {
    if (!b) return IBTFO_E_ARG;
    batch_free(b);
    return IBTFO_OK;
}

extern "C" double ibtfo_batch_clock(const ibtfo_batch* b) { return b ? b->now() : 0.0; }

extern "C" int ibtfo_batch_submit(ibtfo_batch* b, const ibtfo_request* r, long long* ticket) {
    if (!b || !r) { set_error("null argument"); return IBTFO_E_ARG; }
    const ibtfo_config& c = r->cfg;
    const ibtfo_batch_config& k = b->bc;
    if (c.scheme != k.scheme || c.riemann != k.riemann || c.limiter != k.limiter || c.boundary != k.boundary ||
        c.time != k.time || c.precision != k.precision || c.quirks != k.quirks) {
        set_error("ibtfo_batch_submit: the request's slot selection differs from the batch's");
        return IBTFO_E_ARG;
    }
    // a per-cell level set is staged in device memory during the geometry
    const long long need = ibtfo_footprint(&c) +
        (r->ls_kind == IBTFO_LS_FIELD
             ? (long long)(c.N_x + 2 * c.num_ghost_cell) * (c.N_y + 2 * c.num_ghost_cell) * (long long)sizeof(double)
             : 0);
    if (need > (long long)b->arena.capacity()) {
        set_error("ibtfo_batch_submit: the solve needs %lld bytes, more than the batch's budget", need);
        return IBTFO_E_NOMEM;
    }
    if (c.N_x > 0 && c.N_y > 0 &&
        (cfg_step_blocks(c, k.rows) > k.max_step_blocks || cfg_ring_blocks(c) > k.max_ring_blocks)) {
        set_error("ibtfo_batch_submit: the solve needs %d step blocks and %d ring blocks, more than the batch's "
                  "capacity (%d, %d)", cfg_step_blocks(c, k.rows), cfg_ring_blocks(c), k.max_step_blocks,
                  k.max_ring_blocks);
        return IBTFO_E_NOMEM;
    }
    if ((r->ls_nprm > 0 && !r->ls_prm) || (r->init_nprm > 0 && !r->init_prm) ||
        (r->ls_kind == IBTFO_LS_FIELD && !r->phi)) {
        set_error("ibtfo_batch_submit: missing parameter array");
        return IBTFO_E_ARG;
    }
    Req q;
    q.ticket = b->next_ticket++;
    q.cfg = c;
    q.ls_kind = r->ls_kind;
    q.init_kind = r->init_kind;
    q.keep = r->keep_state;
    if (r->ls_nprm > 0) q.ls_prm.assign(r->ls_prm, r->ls_prm + r->ls_nprm);
    if (r->init_nprm > 0) q.init_prm.assign(r->init_prm, r->init_prm + r->init_nprm);
    if (r->ls_kind == IBTFO_LS_FIELD) {
        const size_t n = (size_t)(c.N_x + 2 * c.num_ghost_cell) * (c.N_y + 2 * c.num_ghost_cell);
        q.phi.assign(r->phi, r->phi + n);
    }
    q.arrival = r->arrival < 0 ? b->now() : r->arrival;
    b->queue.push_back(std::move(q));
    if (ticket) *ticket = b->queue.back().ticket;
    return IBTFO_OK;
}

// Retires the queued request at the head with status rc (it never ran).
static void retire_head(ibtfo_batch* b, int rc) {
    Req& q = b->queue.front();
    Done d{};
    d.r.status = rc;
    d.r.t_arrival = q.arrival;
    d.r.t_admit = d.r.t_done = b->now();
    b->done[q.ticket] = std::move(d);
    b->ready.push_back(q.ticket);
    b->finished++;
    b->queue.pop_front();
}

// Builds the queued request at the head into a context on the admission stream.
// Returns 1 if admitted, 0 if it must wait (memory or grid capacity), <0 on a
// request error (the request is then retired with that status).
static int admit_one(ibtfo_batch* b, int reserved_step, int reserved_ghost, int reserved_ring) {
    Req& q = b->queue.front();
    ibtfo_ctx* x = nullptr;
    int rc = ctx_create(&q.cfg, &b->arena, b->adm, b->bc.rows, &x);
    if (rc == IBTFO_E_NOMEM) return 0;
    if (rc == IBTFO_OK)
        rc = ibtfo_geometry(x, q.ls_kind, q.ls_prm.empty() ? nullptr : q.ls_prm.data(), (int)q.ls_prm.size(),
                            q.phi.empty() ? nullptr : q.phi.data());
    if (rc == IBTFO_E_NOMEM) {  // the ghost list did not fit the arena: wait for memory
        free_ctx(x);
        return 0;
    }
    if (rc == IBTFO_OK)
        rc = ibtfo_initialize(x, q.init_kind, q.init_prm.empty() ? nullptr : q.init_prm.data(),
                              (int)q.init_prm.size());
    if (rc == IBTFO_OK) rc = ibtfo_prepare(x, q.cfg.boundary != IBTFO_REGULAR);
    if (rc == IBTFO_OK || rc == IBTFO_E_NONFINITE) {
        if (b->used_step + reserved_step + step_blocks(x) > b->bc.max_step_blocks ||
            b->used_ghost + reserved_ghost + ghost_blocks(x) > b->bc.max_ghost_blocks ||
            b->used_ring + reserved_ring + ring_blocks(x) > b->bc.max_ring_blocks) {
            free_ctx(x);
            return 0;
        }
        b->pending.emplace_back(x, std::move(q));
        b->pending_t.push_back(b->now());
        b->queue.pop_front();
        return 1;
    }
    if (x) free_ctx(x);
    retire_head(b, rc);
    return -1;
}

static int rebuild_tables(ibtfo_batch* b) {
    const int S = b->bc.max_slots;
    int ns = 0, ngh = 0, nr = 0;
    for (int i = 0; i < S; i++) {
        ibtfo_ctx* x = b->slots[i].ctx;
        if (!x) {
            memset(&b->h_step[i], 0, sizeof(StepArgs));
            memset(&b->h_fin[i], 0, sizeof(FinishArgs));
            memset(&b->h_fill[i], 0, sizeof(FillArgs));
            memset(&b->h_edge[i], 0, sizeof(EdgeArgs));
            continue;
        }
        b->h_step[i] = x->sa;
        b->h_fin[i] = x->fin;
        b->h_fill[i] = x->fa;
        b->h_edge[i] = x->ea;
        // the batch runs every phase of a step as its own launch (the edge ring
        // included, and the RK2 stage flip), whatever a lone solve would fuse
        if (b->h_step[i].edge_fused == 2) b->h_step[i].edge_fused = 0;
        b->h_fin[i].keep_cur = 0;
        for (int lb = 0; lb < step_blocks(x); lb++) b->h_smap[ns++] = make_int2(i, lb);
        for (int gb = 0; gb < ghost_blocks(x); gb++) b->h_gmap[ngh++] = make_int2(i, gb * FILL_THREADS);
        for (int rb = 0; rb < ring_blocks(x); rb++) b->h_rmap[nr++] = make_int2(i, rb * EDGE_THREADS);
    }
    for (int k = ns; k < b->bc.max_step_blocks; k++) b->h_smap[k] = make_int2(-1, 0);
    for (int k = ngh; k < b->bc.max_ghost_blocks; k++) b->h_gmap[k] = make_int2(-1, 0);
    for (int k = nr; k < b->bc.max_ring_blocks; k++) b->h_rmap[k] = make_int2(-1, 0);
    b->used_step = ns;
    b->used_ghost = ngh;
    b->used_ring = nr;
    hipStream_t s = b->run;
    if (b->bc.persistent) {
        b->h_counts[0] = ns; b->h_counts[1] = ngh; b->h_counts[2] = nr;
        HIPCHK(hipMemcpyAsync(b->d_counts, b->h_counts, 3 * sizeof(int), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(b->d_step, b->h_step, S * sizeof(StepArgs), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_fin, b->h_fin, S * sizeof(FinishArgs), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_fill, b->h_fill, S * sizeof(FillArgs), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_edge, b->h_edge, S * sizeof(EdgeArgs), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_smap, b->h_smap, b->bc.max_step_blocks * sizeof(int2), hipMemcpyHostToDevice, s));
    if (b->bc.max_ghost_blocks > 0)
        HIPCHK(hipMemcpyAsync(b->d_gmap, b->h_gmap, b->bc.max_ghost_blocks * sizeof(int2), hipMemcpyHostToDevice, s));
    if (b->bc.max_ring_blocks > 0)
        HIPCHK(hipMemcpyAsync(b->d_rmap, b->h_rmap, b->bc.max_ring_blocks * sizeof(int2), hipMemcpyHostToDevice, s));
    b->dirty = false;
    return IBTFO_OK;
}

// The batched form of one step of the reference loop (the same sequence as a
// single solve's, each launch covering every resident solve).
static int enqueue_batch_step(ibtfo_batch* b) {
    const ibtfo_batch_config& k = b->bc;
    hipStream_t s = b->run;
    auto bounds = [&]() -> int {
        if (k.boundary != IBTFO_REGULAR && k.max_ghost_blocks > 0)
            HIPCHK(launch_fill_batch(b->d_fill, b->d_gmap, k.max_ghost_blocks, s));
        if (k.max_ring_blocks > 0) HIPCHK(launch_edge_batch(b->d_edge, b->d_rmap, k.max_ring_blocks, s));
        if (k.boundary != IBTFO_REGULAR && k.max_ghost_blocks > 0)
            HIPCHK(launch_scatter_batch(b->d_fill, b->d_gmap, k.max_ghost_blocks, s));
        return IBTFO_OK;
    };
    int rc;
    if (k.time == IBTFO_RK2) {
        HIPCHK(launch_step_batch(k.scheme, k.riemann, k.limiter, k.precision, b->d_step, b->d_smap,
                                 k.max_step_blocks, 1, s));
        HIPCHK(launch_finish_batch(b->d_fin, k.max_slots, 2, s));
        if ((rc = bounds())) return rc;
        HIPCHK(launch_step_batch(k.scheme, k.riemann, k.limiter, k.precision, b->d_step, b->d_smap,
                                 k.max_step_blocks, 2, s));
    } else {
        HIPCHK(launch_step_batch(k.scheme, k.riemann, k.limiter, k.precision, b->d_step, b->d_smap,
                                 k.max_step_blocks, 0, s));
    }
    HIPCHK(launch_finish_batch(b->d_fin, k.max_slots, 1, s));
    return bounds();
}

// The chunk as one persistent launch: every resident solve's steps, phases
// separated by grid barriers (see k_persist).
static int enqueue_persistent_chunk(ibtfo_batch* b) {
    const ibtfo_batch_config& k = b->bc;
    PersistArgs p;
    memset(&p, 0, sizeof p);
    p.step = b->d_step; p.fin = b->d_fin; p.fill = b->d_fill; p.edge = b->d_edge;
    p.smap = b->d_smap; p.gmap = b->d_gmap; p.rmap = b->d_rmap; p.counts = b->d_counts;
    p.nslots = k.max_slots;
    p.nsteps = k.chunk;
    p.rk2 = k.time == IBTFO_RK2;
    p.has_fill = k.boundary != IBTFO_REGULAR;
    p.bar = b->d_bar;
    p.err = b->d_err;
    p.spin_limit = 1ll << 22;  // about 4 s of polling before the launch gives up
    HIPCHK(launch_persist(k.scheme, k.riemann, k.limiter, k.precision, p, b->persist_grid, b->run));
    return IBTFO_OK;
}

static int capture_graph(ibtfo_batch* b) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->run, hipStreamCaptureModeThreadLocal));
    int rc = IBTFO_OK;
    if (b->bc.persistent) rc = enqueue_persistent_chunk(b);
    for (int i = 0; !b->bc.persistent && i < b->bc.chunk && rc == IBTFO_OK; i++) rc = enqueue_batch_step(b);
    hipError_t e = hipStreamEndCapture(b->run, &g);
    if (e == hipSuccess && rc == IBTFO_OK && g) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    HIPCHK(e);
    return IBTFO_OK;
}

static int harvest(ibtfo_batch* b) {
    const int S = b->bc.max_slots;
    HIPCHK(launch_gather(b->d_fin, S, b->d_stat, b->run));
    HIPCHK(hipMemcpyAsync(b->h_stat, b->d_stat, S * sizeof(SlotStatus), hipMemcpyDeviceToHost, b->run));
    int perr = 0;
    if (b->bc.persistent) HIPCHK(hipMemcpyAsync(&perr, b->d_err, sizeof(int), hipMemcpyDeviceToHost, b->run));
    HIPCHK(hipStreamSynchronize(b->run));
    if (perr) {
        set_error("persistent step kernel: a grid barrier timed out (blocks not resident together)");
        return IBTFO_E_HIP;
    }
    for (int i = 0; i < S; i++) {
        Slot& sl = b->slots[i];
        if (!sl.ctx || !b->h_stat[i].done) continue;
        ibtfo_ctx* x = sl.ctx;
        Done d{};
        d.r.status = b->h_stat[i].status ? IBTFO_E_NONFINITE : IBTFO_OK;
        d.r.iter = b->h_stat[i].iter;
        d.r.now_t = b->h_stat[i].now_t;
        d.r.dt = b->h_stat[i].dt;
        d.r.mass = b->h_stat[i].mass;
        d.r.M = x->L.M;
        d.r.N = x->L.N;
        d.r.t_arrival = sl.req.arrival;
        d.r.t_admit = sl.t_admit;
        if (sl.req.keep) {
            d.U.resize((size_t)x->L.M * x->L.N * 4);
            int rc = ibtfo_download_state(x, d.U.data());
            if (rc && rc != IBTFO_E_NONFINITE) return rc;
        }
        d.r.t_done = b->now();
        b->cell_updates += (long long)x->L.nx * x->L.ny * d.r.iter;
        b->done[sl.req.ticket] = std::move(d);
        b->ready.push_back(sl.req.ticket);
        b->finished++;
        free_ctx(x);
        sl.ctx = nullptr;
        b->dirty = true;
    }
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_progress(ibtfo_batch* b, int* active, int* queued, int* finished) {
    if (!b) return IBTFO_E_ARG;
    const double t_enter = b->now();
    int rc;
    int nact = 0;
    for (auto& s : b->slots) nact += s.ctx != nullptr;
    if (nact == 0) b->used_step = b->used_ghost = b->used_ring = 0;  // tables not yet rebuilt after the last retirements
    bool launched = false;
    if (nact > 0) {
        if (b->dirty && (rc = rebuild_tables(b))) return rc;
        if (!b->graph && (rc = capture_graph(b))) return rc;
        HIPCHK(hipGraphLaunch(b->graph, b->run));
        launched = true;
    }
    // admissions overlap the chunk (geometry and initial state on the admission
    // stream): they go on while the chunk runs, at least one per call, so that
    // the GPU never waits for a whole wave of set-ups; with no chunk running,
    // a short burst is admitted and the next call starts the GPU
    int free_slots = 0;
    for (auto& s : b->slots) free_slots += s.ctx == nullptr;
    free_slots -= (int)b->pending.size();
    const double tnow = b->now();
    int tried = 0;
    while (!b->queue.empty() && free_slots > 0 && b->queue.front().arrival <= tnow) {
        if (launched ? (tried > 0 && hipStreamQuery(b->run) != hipErrorNotReady) : tried >= ADMIT_BURST) break;
        int rs = 0, rg = 0, rr = 0;
        for (auto& p : b->pending) { rs += step_blocks(p.first); rg += ghost_blocks(p.first); rr += ring_blocks(p.first); }
        int a = admit_one(b, rs, rg, rr);
        if (a == 0 && nact == 0 && b->pending.empty()) {
            // the batch is empty and the request still does not fit (its ghost
            // list exceeds the ghost-block capacity, or the arena): it never will
            set_error("ibtfo_batch_progress: request %lld does not fit an empty batch", b->queue.front().ticket);
            retire_head(b, IBTFO_E_NOMEM);
            tried++;
            continue;
        }
        if (a == 0) break;
        if (a == 1) free_slots--;
        tried++;
    }
    if (launched) {
        HIPCHK(hipStreamSynchronize(b->run));
        b->chunks++;
        b->steps += b->bc.chunk;
        if ((rc = harvest(b))) return rc;
    }
    // admitted solves join the tables for the next chunk
    for (size_t p = 0; p < b->pending.size(); p++) {
        for (int i = 0; i < (int)b->slots.size(); i++) {
            if (b->slots[i].ctx) continue;
            b->slots[i].ctx = b->pending[p].first;
            b->slots[i].req = std::move(b->pending[p].second);
            b->slots[i].t_admit = b->pending_t[p];
            b->dirty = true;
            break;
        }
    }
    b->pending.clear();
    b->pending_t.clear();
    b->busy += b->now() - t_enter;
    int na = 0;
    for (auto& s : b->slots) na += s.ctx != nullptr;
    if (active) *active = na;
    if (queued) *queued = (int)b->queue.size();
    if (finished) *finished = b->finished;
    return IBTFO_OK;
}

// Allocates the persistent kernel's state (item counts, barrier counter, error
// flag) and sizes its grid, once.
static int ensure_persistent(ibtfo_batch* b) {
    if (b->d_counts) return IBTFO_OK;
    const ibtfo_batch_config& k = b->bc;
    b->persist_grid = persist_blocks(k.scheme, k.riemann, k.limiter, k.precision);
    if (b->persist_grid <= 0) { set_error("no resident grid for the persistent step kernel"); return IBTFO_E_HIP; }
    if (dev_alloc(&b->d_counts, 3) || dev_alloc(&b->d_bar, 1) || dev_alloc(&b->d_err, 1) ||
        host_alloc(&b->h_counts, 3) || hipMemset(b->d_err, 0, sizeof(int)) != hipSuccess) {
        set_error("allocation of the persistent kernel's state failed");
        return IBTFO_E_NOMEM;
    }
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_autotune(ibtfo_batch* b, const ibtfo_request* sample, int n, int probe_chunks,
                                    int* best_rows, int* best_persistent, double* best_ms_per_step) {
    if (!b || !sample || n <= 0 || probe_chunks <= 0) { set_error("ibtfo_batch_autotune: bad argument"); return IBTFO_E_ARG; }
    for (auto& s : b->slots)
        if (s.ctx) { set_error("ibtfo_batch_autotune: the batch has resident solves"); return IBTFO_E_STATE; }
    if (!b->queue.empty() || !b->pending.empty()) {
        set_error("ibtfo_batch_autotune: the batch has queued requests");
        return IBTFO_E_STATE;
    }
    const ibtfo_batch_config keep = b->bc;
    static const int rows_cand[] = {4, 8, 16, 32};
    int rc = IBTFO_OK, brow = keep.rows, bper = keep.persistent;
    // candidates are ranked by cell-updates/s: a shorter strip needs more step
    // blocks per solve, so fewer of the sample fit the grid capacity and the
    // batched step has less work
    double bms = 1e300, brate = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("ibtfo_batch_autotune: event creation failed");
        rc = IBTFO_E_HIP;
    }
    auto clear = [&]() {
        (void)hipStreamSynchronize(b->run);
        (void)hipStreamSynchronize(b->adm);
        for (auto& s : b->slots)
            if (s.ctx) { free_ctx(s.ctx); s.ctx = nullptr; }
        if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
        b->dirty = true;
        b->used_step = b->used_ghost = b->used_ring = 0;
    };
    for (int per = 0; per < 2 && !rc; per++) {
        if (per && (rc = ensure_persistent(b))) { rc = IBTFO_OK; break; }  // no persistent grid: launches only
        for (int r : rows_cand) {
            b->bc.rows = r;
            b->bc.persistent = per;
            // the sample as resident solves that run for the whole probe
            int ns = 0, ng = 0, nr = 0, used = 0;
            double cells = 0.0;
            for (int i = 0; i < n && used < b->bc.max_slots && !rc; i++) {
                ibtfo_config c = sample[i].cfg;
                c.max_iter = 1 << 30;
                c.Psy_time = 1e300;
                ibtfo_ctx* x = nullptr;
                int e = ctx_create(&c, &b->arena, b->adm, r, &x);
                if (e == IBTFO_OK)
                    e = ibtfo_geometry(x, sample[i].ls_kind, sample[i].ls_prm, sample[i].ls_nprm, sample[i].phi);
                if (e == IBTFO_OK)
                    e = ibtfo_initialize(x, sample[i].init_kind, sample[i].init_prm, sample[i].init_nprm);
                if (e == IBTFO_OK) e = ibtfo_prepare(x, c.boundary != IBTFO_REGULAR);
                if (e == IBTFO_OK && (ns + step_blocks(x) > b->bc.max_step_blocks ||
                                      ng + ghost_blocks(x) > b->bc.max_ghost_blocks ||
                                      nr + ring_blocks(x) > b->bc.max_ring_blocks))
                    e = IBTFO_E_NOMEM;
                if (e != IBTFO_OK) {  // does not fit (memory or grid capacity): probe with what does
                    if (x) free_ctx(x);
                    if (e != IBTFO_E_NOMEM) rc = e;
                    break;
                }
                ns += step_blocks(x); ng += ghost_blocks(x); nr += ring_blocks(x);
                cells += (double)c.N_x * c.N_y;
                b->slots[used++].ctx = x;
            }
            if (!rc && used == 0) { set_error("ibtfo_batch_autotune: no sample request fits the batch"); rc = IBTFO_E_NOMEM; }
            float ms = 0;
            if (!rc) rc = rebuild_tables(b);
            if (!rc) rc = capture_graph(b);
            if (!rc && (hipStreamSynchronize(b->adm) != hipSuccess || hipGraphLaunch(b->graph, b->run) != hipSuccess ||
                        hipEventRecord(e0, b->run) != hipSuccess)) rc = IBTFO_E_HIP;
            for (int q = 0; q < probe_chunks && !rc; q++)
                if (hipGraphLaunch(b->graph, b->run) != hipSuccess) rc = IBTFO_E_HIP;
            if (!rc && (hipEventRecord(e1, b->run) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) rc = IBTFO_E_HIP;
            int perr = 0;
            if (!rc && per && hipMemcpy(&perr, b->d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = IBTFO_E_HIP;
            if (!rc && perr) { set_error("ibtfo_batch_autotune: a grid barrier of the persistent kernel timed out"); rc = IBTFO_E_HIP; }
            clear();
            if (rc == IBTFO_E_HIP) set_error("ibtfo_batch_autotune: probe failed (%s)", hipGetErrorString(hipGetLastError()));
            if (rc) break;
            const double per_step = ms / (probe_chunks * (double)b->bc.chunk);
            const double rate = cells / per_step;
            if (rate > brate) { brate = rate; bms = per_step; brow = r; bper = per; }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    b->bc = keep;
    if (rc) return rc;
    b->bc.rows = brow;
    b->bc.persistent = bper;
    if (bper && (rc = ensure_persistent(b))) return rc;
    if (best_rows) *best_rows = brow;
    if (best_persistent) *best_persistent = bper;
    if (best_ms_per_step) *best_ms_per_step = bms;
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_poll(ibtfo_batch* b, long long* ticket) {
    if (!b || !ticket) return IBTFO_E_ARG;
    if (b->ready.empty()) return 0;
    *ticket = b->ready.front();
    b->ready.pop_front();
    return 1;
}

extern "C" int ibtfo_batch_result(ibtfo_batch* b, long long ticket, ibtfo_result* res, double* U_OLD) {
    if (!b) return IBTFO_E_ARG;
    auto it = b->done.find(ticket);
    if (it == b->done.end()) { set_error("ibtfo_batch_result: ticket %lld has not finished", ticket); return IBTFO_E_STATE; }
    if (res) *res = it->second.r;
    if (U_OLD) {
        if (it->second.U.empty()) { set_error("ibtfo_batch_result: ticket %lld kept no state", ticket); return IBTFO_E_STATE; }
        memcpy(U_OLD, it->second.U.data(), it->second.U.size() * sizeof(double));
    }
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_release(ibtfo_batch* b, long long ticket) {
    if (!b) return IBTFO_E_ARG;
    b->done.erase(ticket);
    return IBTFO_OK;
}

extern "C" int ibtfo_batch_stats_get(ibtfo_batch* b, ibtfo_batch_stats* st) {
    if (!b || !st) return IBTFO_E_ARG;
    memset(st, 0, sizeof *st);
    long long live = 0;
    for (auto& s : b->slots) {
        if (!s.ctx) continue;
        st->active++;
    }
    (void)live;
    st->queued = (int)b->queue.size();
    st->finished = b->finished;
    st->chunks = b->chunks;
    st->steps = b->steps;
    st->cell_updates = b->cell_updates;
    st->arena_used = (long long)b->arena.used();
    st->arena_capacity = (long long)b->arena.capacity();
    st->busy_s = b->busy;
    return IBTFO_OK;
}
