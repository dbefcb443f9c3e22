// Internal layout and kernel interfaces of the IBTFO HIP library.
//
// Device layout (all FP64 state in SoA planes, x fastest):
//   cell (i, j), 0 <= i < M = N_x + 2ng, 0 <= j < N = N_y + 2ng, lives at
//   plane[off + j*pitch + i]; `off` puts the first interior column (i = ng) on a
//   128-byte boundary and `pitch` is a multiple of 16 doubles, so every row's
//   interior starts cache-line aligned. A state buffer is 4 planes (rho, rho u,
//   rho v, E) spaced `plane` doubles apart; two buffers ping-pong between steps.
//   Cell types are a uint8 plane with the same indexing.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/ibtfo.h"
#include "ibtfo_math.h"

namespace ibtfo {

constexpr int LOGCAP = 4096;  // per-step log ring (dt, time, mass)
constexpr int STRIP_W = 60;   // interior columns updated by one 64-lane block (lanes 2..61)
constexpr int ROWPAR_MAX_ROWS = 7;  // rows of a row-parallel strip (blocks of at most 8 waves)

struct Layout {
    int nx, ny, ng, M, N;
    int pitch, off;
    long long plane;  // doubles between consecutive planes of one buffer
    __host__ __device__ long long at(int i, int j) const { return (long long)off + (long long)j * pitch + i; }
};

// The clock a step kernel runs with, in the deferred-finish form of small
// grids (see StepArgs::df): kernel k of a batch of steps computes it from the
// previous kernel's head and partial records and leaves it in Ctl::df[k & 1]
// for kernel k + 1; `pending` = 1 when kernel k ran its step, whose finish
// (dt, clock, log) the next kernel, or the batch's closing flush, applies.
struct DfHead {
    double dt, now_t, mass;
    int iter, done, cur, pending;
};

// Device control block of one solve: the time loop runs on the device so that
// steps can be queued (or replayed from a graph) without host round trips.
struct Ctl {
    double dt;      // dt for the next step (ComputeDt of the current state)
    double now_t;   // physical time as the reference accumulates it
    int iter;       // steps taken
    int done;       // 1 once now_t >= Psy_time or iter >= max_iter
    int cur;        // which state buffer holds the current state
    int ran;        // 1 if the last launched step executed
    int status;     // 0 ok; 1 non-finite value met in the dt reduction
    unsigned int folded;  // fold blocks done in the running finish (0 between launches)
    double mass_now;                      // CaculateMass of the state seen by the last reduction
    DfHead df[2];                         // deferred finish: the heads of a batch's kernels, by parity
    double log_dt[LOGCAP], log_t[LOGCAP], log_mass[LOGCAP];
};

// Cut-cell record (cut cells only; no reference counterpart): volume fraction,
// face apertures W, E, S, N, wall vector (the closure of the faces) and the
// merge target of a small cell (-1: none).
struct CutCell {
    double alpha, aw, ae, as, an, sbx, sby;
    long long target;
};
// A merge group: a target cell and the small cut cells that chose it
// (neighbours W, E, S, N in that order), with their volume fractions.
struct MergeGroup {
    long long target;
    long long member[4];
    double at;
    double am[4];
    int n, pad;
};

struct StepArgs {
    double* buf[2];     // state buffers
    Layout L;
    const uint8_t* type;  // cell types; the cut variant's plane adds MERGED_BIT
    Ctl* ctl;
    double* mass_part;  // per-block partials, PART_STRIDE doubles each
    const double* corner;  // [2 sides][2ng rows][4]: U_NEW ring at columns ng / N_x+ng-1
    double gam, gm1, gp1, c1, dx, dy, cfl, t_end;
    double r_gam, r_gm1, r_dx, r_dy;  // correctly rounded reciprocals of the constant divisors
    int max_iter;
    int quirks;
    int edge_fused;         // the step kernel writes the output's ring: 1 the reference
                            // Global_Boundary ring, 2 the periodic ring (Regular boundary,
                            // nx, ny >= ng); 0: an edge kernel writes it after the step
    // SimpleBound: the reference's U_NEW of non-fluid cells, which its Boundary
    // never refreshes (no copy back), as Global_Boundary reads it; else null
    const double* unew_nf;
    int corner_from_input;  // corners from the input buffer's ring (GFM) or `corner` (Regular)
    int rows;               // rows per block strip
    int nbx, nby;           // grid
    int stage;              // 0: single-step update; RK2 (Heun): 1 = U1 = U^n + dt L(U^n),
                            // 2 = (U^n + U1 + dt L(U1)) / 2, written over U^n in place
                            // (after a finish made U1 current); 3 = stage 2 with no such
                            // flip: reads U1 from buf[cur ^ 1], writes over U^n in buf[cur]
    const CutCell* cut;     // cut-cell records (type-3 cells), else null
    const int* cut_index;   // plane: record of each type-3 cell
    int rowpar;             // row-parallel form (k_step_rowpar: rows + 1 waves per block)
    // Deferred finish (row-parallel form, single-step, Regular boundary with a
    // ring the step kernel writes, at most DF_MAX_REC blocks): no finish kernel
    // between steps. Kernel k of a batch reduces kernel k-1's partial records
    // (df_part[(k-1) & 1], in the one-wave finish's order) and advances the
    // clock itself, every block alike; block (0,0) keeps the log. The batch
    // ends with k_df_flush, which publishes the last head to the control block.
    int df_ok;              // the solve qualifies
    int df;                 // this launch: 0 off; 1 first kernel of a batch (clock from Ctl); 2 later
    int df_k;               // this launch's parity in its batch (k & 1)
    double* df_part;        // 2 x DF_MAX_REC partial records
};
constexpr int DF_MAX_REC = 256;  // blocks of a deferred-finish step (4 records per lane of one wave)

// Small-cell merging after a cut-cell step (one thread per group).
struct MergeArgs {
    double* buf[2];
    long long plane;
    Ctl* ctl;
    const MergeGroup* groups;
    int ngroups;
    double gm1, r_gm1;
    int quirks;
    double dx, dy;
    double* part;  // PART_STRIDE doubles per merge block: dt / mass partials of the groups' cells
};
constexpr int MERGE_THREADS = 64;
// in the cut-cell step kernel's type plane: a merge group's cell (target or
// member), whose dt / mass partials the merge kernel writes
constexpr uint8_t MERGED_BIT = 4;
inline int merge_blocks(int ngroups) { return (ngroups + MERGE_THREADS - 1) / MERGE_THREADS; }

// Ghost-cell records (compacted list of interior cells of type 2), built once.
struct Ghost {
    long long cell;     // index of the ghost cell in a plane
    long long d1;       // index of donor (IDX, IDY) of the image point; donors 2..4 = +pitch, +pitch+1, +1
    long long d1e;      // same for the extra point (BCGFM/HBCGFM)
    double nx, ny;      // level-set normal at the ghost
    double phi;         // |phi| at the ghost
    double a[4];        // image-point weights (IDW/BCGFM); Linear/RiemannBaseGFM: a[0] = chosen donor
    double ae[4];       // extra-point weights
    double xm, ym, xe, ye;  // image (mirror) and extra points
    int idx, idy, idxe, idye, pick;  // SimpleBound: idx/idy = donor offsets IDX/IDY, idxe/idye = ghost
                                     // index of an x/y donor that is a ghost filled earlier (else -1),
                                     // pick = dependency level
    int ring;           // bit q (image) / 4+q (extra): donor q lies in the domain ring
};

struct FillArgs {
    double* buf[2];
    Layout L;
    Ctl* ctl;
    const Ghost* ghosts;
    int nghost;
    double* out;  // [nghost][4]
    int bc;
    double gam, gm1;
    double dxg, dyg;   // spacing from the coordinates (BCGFM delta)
    int use_cur;       // 1: operate on ctl->cur (after a step); 0: on buffer `fixed`
    int fixed;
    int require_ran;   // skip unless the preceding step executed
    int ring_prev;     // read ring donors from the previous state (the ring before Global_Boundary)
    int quirks;
    int level;         // SimpleBound: fill only the ghosts of this dependency level (Ghost::pick)
};

struct EdgeArgs {
    double* buf[2];
    Layout L;
    Ctl* ctl;
    int kind;
    double gam;
    double inflow[4];  // conserved inflow state
    int require_ran;
    // SimpleBound: non-fluid source cells are read from the U_NEW its Boundary
    // leaves unrefreshed (as the edge policies read U_NEW); else null
    const double* unew_nf;
    const uint8_t* type;
};

struct ReduceArgs {
    const double* U;  // state buffer
    const CutCell* cut;  // cut cells count with their volume fraction
    const int* cut_index;
    Layout L;
    const uint8_t* type;
    Ctl* ctl;
    double* mass_part;
    double gam, gm1, dx, dy, cfl, t_end;
    int max_iter;
    int nblocks;
    int reset;  // 1: set iter = 0, now_t = 0 (start of a run)
    // when cur_ctl is set, the state is ubuf[cur_ctl->cur], chosen on the
    // device (no host read of the current buffer); U is then unused
    const double* ubuf[2];
    const Ctl* cur_ctl;
};

// the dt / mass reduction over the step kernel's per-block partials
constexpr int PART_STRIDE = 8;  // doubles per block: max a^2, max |u|, max |v|, mass, flag, pad
constexpr int FINISH_THREADS = 1024;
constexpr int FINISH_SMALL_MAX = 512;  // at most this many partial records: the one-wave finish
struct FinishArgs {
    Ctl* ctl;
    const double* part;
    int nblocks;
    double cfl, dx, t_end;
    int max_iter;
    int advance;  // 1: a step was taken (advance the clock); 0: stand-alone ComputeDt;
                  // 2: RK2 stage boundary (make the stage-1 buffer current, no reduction)
    int reset;    // with advance = 0: start the run clock
    int keep_cur; // with advance = 1: leave cur as it is (an RK2 step whose stage 2 ran as
                  // stage 3, over U^n in the current buffer), else make the output current
    // two-level reduction (advance = 1 with many partials): nfold blocks each fold
    // `fold_per` consecutive partials into fold[] first, then one block finishes
    double* fold;
    int nfold, fold_per;
    // a second run of partials after part (the merge kernel's), null if none
    const double* part2 = nullptr;
    int nblocks2 = 0;
};
constexpr int FOLD_THREADS = 256;
constexpr int FOLD_MAX = 128;      // fold records
constexpr int FOLD_MIN_PARTS = 2048;  // fewer partials: one block reads them all

constexpr int FILL_THREADS = 128;  // ghost-fill and scatter blocks
constexpr int EDGE_THREADS = 256;  // edge-ring blocks

// Persistent batched step (k_persist): one launch runs `nsteps` whole steps of
// every resident solve of a batch, its phases (strips; per-solve finish; ghost
// fill and edge ring; scatter) separated by grid barriers among the launch's
// blocks, which are all resident at once (the grid is sized to the kernel's
// occupancy). The item counts are read from device memory, so a captured launch
// stays valid while solves come and go.
struct PersistArgs {
    const StepArgs* step;
    const FinishArgs* fin;
    const FillArgs* fill;
    const EdgeArgs* edge;
    const int2* smap;   // strip items {slot, strip}
    const int2* gmap;   // ghost items {slot, first ghost} (FILL_THREADS ghosts each)
    const int2* rmap;   // ring items {slot, first ring cell} (EDGE_THREADS cells each)
    const int* counts;  // device: {strip items, ghost items, ring items}
    int nslots;
    int nsteps;
    int rk2;
    int has_fill;       // a ghost-fluid boundary (fill + scatter phases)
    unsigned int* bar;  // barrier arrivals (zeroed before each launch)
    int* err;           // set when a barrier wait exceeds spin_limit
    long long spin_limit;
};

// per-slot status record of the batch engine (k_gather)
struct SlotStatus {
    int iter, done, status, pad;
    double now_t, dt, mass;
};

// kernel launchers (ibtfo_step.hip)
hipError_t launch_step(int scheme, int rs, int lim, int prec, const StepArgs& a, hipStream_t s);
hipError_t launch_fill(const FillArgs& a, hipStream_t s);
hipError_t launch_scatter(const FillArgs& a, hipStream_t s);
hipError_t launch_edge(const EdgeArgs& a, hipStream_t s);
hipError_t launch_fill_edge(const FillArgs& f, const EdgeArgs& e, hipStream_t s);  // fill phase 1 + ring
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s);
hipError_t launch_merge(const MergeArgs& a, hipStream_t s);
// a uniform state (conserved U[0..3]) into every element of s0, s1 and s2 (if
// not null); buffer 0 made current, the step flag cleared
hipError_t launch_uniform_state(double* s0, double* s1, double* s2, long long plane, const double* U, Ctl* ctl,
                                hipStream_t s);
hipError_t launch_finish(const FinishArgs& f, hipStream_t s);
// batched forms: one launch over many solves (maps: block -> {slot, first index})
hipError_t launch_step_batch(int scheme, int rs, int lim, int prec, const StepArgs* slots, const int2* map,
                             int nblocks, int stage, hipStream_t s);
hipError_t launch_finish_batch(const FinishArgs* f, int nslots, int advance, hipStream_t s);
hipError_t launch_fill_batch(const FillArgs* f, const int2* map, int nblocks, hipStream_t s);
hipError_t launch_scatter_batch(const FillArgs* f, const int2* map, int nblocks, hipStream_t s);
hipError_t launch_edge_batch(const EdgeArgs* e, const int2* map, int nblocks, hipStream_t s);
hipError_t launch_gather(const FinishArgs* f, int nslots, SlotStatus* out, hipStream_t s);
// the persistent batched step: a zeroing of the barrier counter, then one launch of
// `blocks` blocks (persist_blocks: the grid the kernel can hold resident at once)
hipError_t launch_persist(int scheme, int rs, int lim, int prec, const PersistArgs& p, int blocks, hipStream_t s);
int persist_blocks(int scheme, int rs, int lim, int prec);
hipError_t launch_step_lax(const StepArgs& a, hipStream_t s);
// the row-parallel step kernel (StepArgs::rowpar; 1 <= rows <= ROWPAR_MAX_ROWS), ibtfo_rowpar.hip
hipError_t launch_step_rowpar(int scheme, int rs, int lim, int prec, const StepArgs& a, hipStream_t s);
// the closing kernel of a deferred-finish batch whose last kernel had parity k
hipError_t launch_df_flush(const StepArgs& a, int k, hipStream_t s);
hipError_t launch_edge_ref(const EdgeArgs& a, const double* corner, int corner_from_input, hipStream_t s);

}  // namespace ibtfo
