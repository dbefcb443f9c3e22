/*
 * ibtfo.h -- C ABI of the MI355X-native IBTFO per-timestep Euler update.
 *
 * One opaque context per solve. All state lives in device memory (SoA planes,
 * one FP64 plane per conserved variable, plus a uint8 cell-type plane); the
 * host exchanges it in the reference's AoS layout. Every call returns an int
 * status (IBTFO_OK = 0). No call allocates unless its comment says so.
 *
 * Reference interfaces replaced (reference-relative paths):
 *   ibtfo_config_from_inp   ParamReader::open/numbers/strings   Src/Utils/ParamReader.cpp:14-68
 *                           + Makefile slot selection            Test/Sod/Makefile:1-9
 *   ibtfo_create/destroy    acc enter/exit data create/copyout   Test/Sod/Sod.cpp:74-78, :187-191
 *   ibtfo_upload_state      acc update device                    Test/Sod/Sod.cpp:88-92
 *   ibtfo_download_*        acc update self                      Test/Sod/Sod.cpp:100-104
 *   ibtfo_geometry          Psy_coord + Level_Set(+hook) + Scheme_Index
 *                                       Src/Geometry/Psy_coord.H:11-20, Src/Geometry/Level_Set.H:12-18,
 *                                       Src/Geometry/Level_Set_function.H:13-19,
 *                                       Src/BoundaryCondition/Scheme_Index.H:13-19
 *   ibtfo_initialize        Initialize (user hook)               Src/Base/Common/Initialize.H:12-23
 *   ibtfo_boundary          Boundary                             Src/BoundaryCondition/Boundary.H:11-20
 *   ibtfo_compute_dt        ComputeDt                            Src/Base/Common/ComputeDt.H:12-23
 *   ibtfo_step              TimeAdvance + Boundary + Global_Boundary + ComputeDt (+ CaculateMass)
 *                                       Src/Base/TimeAdvance/TimeAdvance.H:13-26, Src/BoundaryCondition/Boundary.H:11-29,
 *                                       Test/Sod/Sod.cpp:154-177
 *   ibtfo_mass              CaculateMass                         Src/Utils/CaculateMass.H:14-23
 */
#ifndef IBTFO_H_
#define IBTFO_H_

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define IBTFO_OK 0
#define IBTFO_E_ARG 1       /* bad argument or unsupported combination */
#define IBTFO_E_HIP 2       /* HIP runtime error */
#define IBTFO_E_NOMEM 3     /* device or host allocation failed */
#define IBTFO_E_IO 4        /* input file could not be read */
#define IBTFO_E_STATE 5     /* call out of order (e.g. step before geometry) */
#define IBTFO_E_NONFINITE 6 /* non-finite or non-positive density/pressure seen */

/* slot values use the reference Makefile's names (Test/Sod/Makefile:1-5) */
enum ibtfo_scheme { IBTFO_LAX = 0, IBTFO_MUSCL = 1, IBTFO_MHM = 2, IBTFO_GODUNOV = 3 };
enum ibtfo_riemann { IBTFO_HLL = 0, IBTFO_HLLC = 1, IBTFO_PVRS = 2, IBTFO_STEGER_WARMING = 3 };
enum ibtfo_limiter { IBTFO_MINMOD = 0, IBTFO_VANLEER = 1, IBTFO_VANALBADA = 2 };
enum ibtfo_boundary {
    IBTFO_REGULAR = 0, IBTFO_IDW_GFM = 1, IBTFO_LINEAR_GFM = 2, IBTFO_RIEMANN_GFM = 3,
    IBTFO_BCGFM = 4, IBTFO_HBCGFM = 5,
    IBTFO_CUTCELL = 6, /* cut cells with small-cell merging (no reference counterpart; corrected mode,
                          single-step, not in batches) */
    IBTFO_SIMPLE_BOUND = 7 /* SimpleBound/Scheme_Index.cpp:5-66 + LS_SimpleBound.cpp:6-110 (single-device
                              signatures in the reference; SCHEME_IDX = 2 doubles {IDX, IDY}; not in batches) */
};
enum ibtfo_time { IBTFO_SINGLESTEP = 0, IBTFO_RK2 = 1 };
/* Global_Boundary policies: the reference hook (Test/Sod/Global_Boudary.cpp:24-75),
 * periodic in x and y, supersonic inflow at x = lo_x with zero-gradient elsewhere. */
enum ibtfo_edge { IBTFO_EDGE_REFERENCE = 0, IBTFO_EDGE_PERIODIC = 1, IBTFO_EDGE_INFLOW = 2,
                  IBTFO_EDGE_PNGGEO = 3 /* Test/PNGGeo/Global_Boudary.cpp:14-79: rows as the Sod hook,
                                           right zero-gradient, left held at edge_state (the .inp
                                           default is the hook's Mach 1.3 post-shock state), ring
                                           corners untouched */ };
enum ibtfo_precision { IBTFO_FP64 = 0, IBTFO_MIXED = 1 };

/* level-set hooks (Level_Set_function user hook) */
enum ibtfo_levelset_kind {
    IBTFO_LS_NONE = 0,     /* phi = 1 everywhere: Test/Sod/Level_Set_function.cpp:26-35 */
    IBTFO_LS_CHANNEL = 1,  /* inclined channel, prm {angle_deg, AxisD}: Test/InclineSod/Level_Set_function.cpp:8-56 */
    IBTFO_LS_ANNULUS = 2,  /* annulus about (0.5,0.5), prm {f0, f1}: Test/Vortex/Level_Set_function.cpp:15-40 */
    IBTFO_LS_FIELD = 3,    /* phi given per cell (host array, reference AoS cell order) */
    IBTFO_LS_ELLIPSE = 4,  /* solid ellipse, prm {xc, yc, a, b, angle_deg}; phi = signed distance estimate */
    IBTFO_LS_SHAPES = 5    /* union of bodies, exact signed distance: prm {n, n x [type (0 circle, 1 convex
                              polygon), cx, cy, r, nverts, x0, y0, ..., x7, y7]} (21 doubles per body,
                              polygon vertices counter-clockwise) */
};

/* initialisers (Initialize user hook) */
enum ibtfo_init_kind {
    IBTFO_INIT_SOD_ROWS = 0,  /* Test/Sod/Initialize.cpp:29-52, prm {rho,u,v,p}_L {rho,u,v,p}_R */
    IBTFO_INIT_INCLINE = 1,   /* Test/InclineSod/Initialize.cpp:21-72, prm {angle_deg, L state, R state} */
    IBTFO_INIT_VORTEX = 2,    /* Test/Vortex/Initialize.cpp:15-98, prm {fa, fw} */
    IBTFO_INIT_SOD_X = 3,     /* Sod split at x = prm[0], prm[1..8] = L, R states */
    IBTFO_INIT_UNIFORM = 4,   /* prm {rho,u,v,p} */
    IBTFO_INIT_SHU_VORTEX = 5, /* isentropic vortex, prm {x0, y0, beta, rho_inf, u_inf, v_inf, p_inf} */
    IBTFO_INIT_SHOCK = 6       /* Test/PNGGeo/Initialize.cpp:14-54: prm {x_s, Ms, rho, u, v, p}; that state
                                  where x > x_s, its Mach-Ms normal-shock image behind (the hook's
                                  x_s = 0.1, Ms = 1.3) */
};

typedef struct ibtfo_config {
    /* .inp numbers (Test/Sod/Sod.cpp:34-48) */
    int N_x, N_y, num_ghost_cell, max_iter;
    double lo_x, lo_y, hi_x, hi_y;
    double Psy_time, CFL_number, gamma;
    /* slot selection */
    int scheme, riemann, limiter, boundary, time, edge, precision;
    /* 1: reproduce the reference arithmetic including its defects (HLL y-branch
     * aliasing, VanLeer macro expansion, PVRS y store); 0: corrected forms */
    int quirks;
    double edge_state[4]; /* IBTFO_EDGE_INFLOW: inflow (rho, u, v, p) */
} ibtfo_config;

typedef struct ibtfo_ctx ibtfo_ctx;

/* Fill *cfg with defaults (the reference driver's Params::get defaults) and then
 * the numbers and slot strings found in an .inp file. */
int ibtfo_config_default(ibtfo_config* cfg);
int ibtfo_config_from_inp(const char* path, ibtfo_config* cfg);
/* Numeric value of one key as the reference's ParamReader::numbers() parses it;
 * returns 1 if present, 0 if missing (then *value is untouched), <0 on I/O error. */
int ibtfo_inp_number(const char* path, const char* key, double* value);
/* String value of one key as ParamReader::strings() returns it (inline comments
 * included); returns length, 0 if missing, <0 on error. */
int ibtfo_inp_string(const char* path, const char* key, char* buf, int cap);

/* Allocates all device memory of the solve (state x2, cell types, geometry,
 * scheme index, ghost list, control block) on the current HIP device. */
int ibtfo_create(const ibtfo_config* cfg, ibtfo_ctx** out);
int ibtfo_destroy(ibtfo_ctx* ctx);
/* Run on this HIP stream (hipStream_t, default priority); NULL = the context's own stream.
 * Waits for the work already queued on the previous stream first. */
int ibtfo_set_stream(ibtfo_ctx* ctx, void* stream);
/* sizes: M = N_x + 2 ng columns, N = N_y + 2 ng rows */
int ibtfo_dims(const ibtfo_ctx* ctx, int* M, int* N);
/* device footprint in bytes of a configuration (before create) */
long long ibtfo_footprint(const ibtfo_config* cfg);

/* Geometry on the device: Psy_coord + level-set hook + normals + classification
 * + Scheme_Index (mirror points, donor indices, weights) + compacted ghost list.
 * phi_host: IBTFO_LS_FIELD only (M*N doubles, reference cell order), else NULL. */
int ibtfo_geometry(ibtfo_ctx* ctx, int ls_kind, const double* prm, int nprm, const double* phi_host);
/* Mask level set, Test/PNGGeo/Level_Set.cpp:11-364 (the reference reads Demo.png
 * from the working directory with libpng, Src/Utils/PngProcess.cpp:4-72): phi =
 * sign * distance to the nearest edge pixel of the mask (+ where R+G+B > 0),
 * computed on the device, then normals, classification and ghost records as
 * ibtfo_geometry. rgba: width x height x 4 bytes, top row first. */
int ibtfo_geometry_mask(ibtfo_ctx* ctx, const unsigned char* rgba, int width, int height);
/* The same from a PNG file (8-bit RGBA after the reference reader's transforms:
 * 16-bit cut to the high byte, palette and gray expanded, filler alpha). */
int ibtfo_geometry_png(ibtfo_ctx* ctx, const char* path);
/* Decodes a PNG file into rgba (cap bytes; NULL to query the size only). */
int ibtfo_png_decode(const char* path, unsigned char* rgba, long long cap, int* width, int* height);

/* Initialize hook (Test/<Case>/Initialize.cpp): the state becomes both U_OLD and
 * U_NEW. IBTFO_INIT_UNIFORM is written on the device, ordered on the solve's
 * stream (no synchronisation); the other kinds are evaluated on the host and
 * uploaded. */
int ibtfo_initialize(ibtfo_ctx* ctx, int init_kind, const double* prm, int nprm);
/* Boundary once (the drivers' call before the loop for GFM variants,
 * Test/InclineSod/InclineSod.cpp:81), then ComputeDt; resets time/iteration. */
int ibtfo_prepare(ibtfo_ctx* ctx, int apply_boundary);

/* state exchange in the reference's AoS layout: U[(M*j+i)*4+k] (M*N*4 doubles).
 * U_NEW may be NULL on upload (then U_NEW = U_OLD, as every Initialize hook does). */
int ibtfo_upload_state(ibtfo_ctx* ctx, const double* U_OLD, const double* U_NEW);
int ibtfo_download_state(ibtfo_ctx* ctx, double* U_OLD);
/* XYCOORD[(M*j+i)*6+c], c = {x, y, phi, n_x, n_y, type}; ring normals are 0 */
int ibtfo_download_geometry(ibtfo_ctx* ctx, double* XYCOORD);
/* SCHEME_IDX in the selected Boundary's record width (1, 4, 8 or 16 doubles/cell) */
int ibtfo_download_scheme_index(ibtfo_ctx* ctx, double* SCHEME_IDX);
int ibtfo_sch_width(const ibtfo_ctx* ctx);
int ibtfo_num_ghosts(const ibtfo_ctx* ctx);
/* Cut-cell boundary: number of cut cells (type 3) and of small-cell merge
 * groups after ibtfo_geometry* (0 for other boundaries). No reference
 * counterpart: the reference has no cut-cell method. */
int ibtfo_cut_counts(const ibtfo_ctx* ctx, int* cut_cells, int* merge_groups);

/* WriteData (Src/Utils/WriteData.cpp:5-247): the reference's output files for
 * the current state, byte for byte -- dat_0_<iter>.vtk (legacy-ASCII VTK
 * STRUCTURED_POINTS: density, x-vel, y-vel, pressure, Energy, Mach_Number,
 * velocity of the interior cells as "%f" lines; -1 / -100 on non-fluid cells)
 * and result.visit, appended to ("!NBLOCKS 1" first when iter == 0). Files go
 * to `dir` (NULL or "" = the working directory). ibtfo_write_vtk downloads the
 * state and geometry of a solve; ibtfo_vtk_write_host formats host arrays in
 * the reference's AoS layout (no device needed). */
int ibtfo_write_vtk(ibtfo_ctx* ctx, const char* dir, int iter, double now_t);
int ibtfo_vtk_write_host(const ibtfo_config* cfg, int iter, double now_t, const double* U_OLD,
                         const double* XYCOORD, const char* dir);

/* Up to nsteps steps of the reference loop body (Test/Sod/Sod.cpp:154-177):
 * TimeAdvance(dt); Boundary; Global_Boundary; dt = ComputeDt; mass; now_t += dt;
 * iter++ -- stopping, on the device, when now_t >= Psy_time or iter >= max_iter.
 * Asynchronous on the context stream; dt never leaves the device. */
int ibtfo_step(ibtfo_ctx* ctx, int nsteps);
/* One step with the caller's dt: TimeAdvance's dt argument (Src/Base/TimeAdvance/TimeAdvance.H:13-26,
 * passed on to Advance by SingleStep.cpp:5-20) followed by Boundary, Global_Boundary and ComputeDt,
 * whatever the device clock's stop test says (the caller keeps its own clock, e.g. a driver that
 * clips the last step to Psy_time). Synchronises before queueing the step. */
int ibtfo_step_dt(ibtfo_ctx* ctx, double dt);
/* Same as ibtfo_step, replayed from a captured HIP graph of `chunk` steps (captured on first use;
 * a new geometry, stream or tile shape drops the captured graph). */
int ibtfo_step_graph(ibtfo_ctx* ctx, int nsteps, int chunk);
/* Synchronises; iteration count, physical time and the current dt. */
int ibtfo_status(ibtfo_ctx* ctx, int* iter, double* now_t, double* dt);
/* Per-step log of the last `count` (<= 4096) steps ending at the current step:
 * dt used after the step, time, fluid mass (CaculateMass). Any pointer may be NULL. */
int ibtfo_log(ibtfo_ctx* ctx, int count, double* dt, double* t, double* mass);
/* ComputeDt on the current state (synchronises). */
int ibtfo_compute_dt(ibtfo_ctx* ctx, double* dt);
/* CaculateMass on the current state (synchronises). */
int ibtfo_mass(ibtfo_ctx* ctx, double* mass);
/* No counterpart in the reference: for callers that cannot allocate or
 * synchronise (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). ibtfo_compute_dt + ibtfo_mass of the current state, queued on the
 * solve's stream: out (device memory, 3 doubles) gets {dt, mass, 1 if a
 * non-physical fluid state was seen else 0}, bit for bit what ibtfo_compute_dt
 * and ibtfo_mass return. Allocates nothing and does not synchronise: the block
 * partials and control record live in a fixed device array of the library
 * (about 360 KB per GPU, held until the process ends), so at most 4096 x 256 =
 * 1,048,576 interior cells (larger solves: IBTFO_E_ARG). Calls that overlap on
 * one GPU share that array and race. Needs ibtfo_prepare first; the current
 * device must be the solve's. */
int ibtfo_compute_dt_noalloc(ibtfo_ctx* ctx, double* out);
/* Boundary (the selected variant) once on the current state. */
int ibtfo_boundary(ibtfo_ctx* ctx);

/* Timing: runs nsteps steps between two HIP events recorded on the context
 * stream (replaying the captured graph when ibtfo_step_graph has captured one)
 * and returns the elapsed milliseconds; *step_kernel_ms gets the mean
 * duration of the fused step kernel alone measured with events around each
 * launch in a second pass of `nsteps` steps (may be NULL to skip). */
int ibtfo_time_steps(ibtfo_ctx* ctx, int nsteps, double* total_ms, double* step_kernel_ms);

/* Step-kernel tile shape: a block updates a strip of 60 columns x `rows` rows.
 * rows = 0 restores the default shape. May allocate (block partials). Same as
 * ibtfo_set_tile(ctx, rows, 0). */
int ibtfo_set_rows(ibtfo_ctx* ctx, int rows);
/* Tile shape and form: row_parallel = 0 is the marching strip (one wave walks
 * the strip's rows one after the other); row_parallel = 1 gives each of the
 * strip's rows (1..7) its own wave, plus one for the face below the strip, so a
 * step's latency is one row's instead of rows + 1 (small grids). Both forms
 * give the same state bit for bit. Not for the Lax scheme. */
int ibtfo_set_tile(ibtfo_ctx* ctx, int rows, int row_parallel);
/* current rows per strip and the step grid (blocks in x and y) */
int ibtfo_tile(const ibtfo_ctx* ctx, int* rows, int* nbx, int* nby);
/* the same, and the form (1: row-parallel) */
int ibtfo_tile_ex(const ibtfo_ctx* ctx, int* rows, int* nbx, int* nby, int* row_parallel);
/* 1 when a step of this solve runs as one kernel with the current tile (the
 * deferred finish: row-parallel strips of a single-step solve on the Regular
 * boundary with the reference edge hook or periodic edges, at most 256 blocks;
 * each ibtfo_step
 * call or graph replay then ends with one closing kernel), 0 when a finish
 * kernel follows each step kernel, -1 for a null context. */
int ibtfo_one_kernel_steps(const ibtfo_ctx* ctx);
/* Tile-shape autotuner: times `probe_steps` steps of the current solve for each
 * candidate shape (marching strips of 1 .. 256 rows, row-parallel strips of 1 ..
 * 7 rows), restoring the state and clock after each probe, and keeps the
 * fastest (ibtfo_tile_ex reports its form). Allocates a temporary snapshot of
 * both state buffers. Fails with IBTFO_E_STATE when the run would end within a
 * probe. */
int ibtfo_autotune(ibtfo_ctx* ctx, int probe_steps, int* best_rows, double* best_ms_per_step);

/* Device pointers of the current state planes (rho, rho u, rho v, E), their
 * row pitch in doubles and the offset of cell (0,0) in each plane. */
int ibtfo_device_state(ibtfo_ctx* ctx, double** planes4, int* pitch, int* offset);

/* ------------------------------------------------------------------------
 * The inner operators on the reference's per-cell scratch record (its U_TMP:
 * M*N cells x 28 doubles, cell (i,j) at rec[(M*j+i)*28 + 4*slot + k];
 * Src/Base/Euler/Advection/MUSCL/EQDefine.H:33-44), for drivers that assemble
 * their own scheme from them. FP64, the step kernel's per-face arithmetic. */
/* Riemann_solverX (dir 0) / Riemann_solverY (dir 1), Src/Base/Euler/Riemann/
 * Riemann_solver.H:22-42, with the solve's Riemann slot, gamma and quirks, on
 * all M*N cells: reads slots 5 and 6 (U_L, U_R / U_D, U_U); HLL and HLLC also
 * write their physical fluxes to slots 3 and 4 (F_L, F_R / G_D, G_U); the face
 * flux goes to slot 2 (F_OLD / G_OLD). Reference mode keeps the reference's y
 * stores: HLL stores nothing when S_D >= 0 (HLL/Riemann_solver.cpp:224-226);
 * PVRS leaves the energy component on the branch of its stray store
 * (PVRS/Riemann_solver.cpp:550, whose out-of-record target is not written).
 * `rec` is device memory; queued on the solve's stream. */
int ibtfo_riemann_record(ibtfo_ctx* ctx, int dir, double* rec);
/* The same on a record in host memory (copied in, solved, copied back before
 * returning). */
int ibtfo_riemann_record_host(ibtfo_ctx* ctx, int dir, double* rec);
/* Conserve2Flux / Conserve2FluxX / Conserve2FluxY (Src/Base/Common/Conserve2Flux.H:11-37):
 * the physical flux of `ncells` conserved states, state c at u[c*stride .. +3],
 * its x flux to f[c*stride ..] and its y flux to g[c*stride ..] (f or g may be
 * NULL; stride >= 4 doubles; the reference calls it with stride 28 on slots
 * of the scratch record). Device memory, queued on `stream` (a hipStream_t;
 * NULL = the null stream). */
int ibtfo_flux_record(double gamma, const double* u, double* f, double* g, long long ncells, int stride,
                      void* stream);
/* The same on host memory (synchronous). Every state is read before any flux
 * is written, so f or g may point into the array u points into. */
int ibtfo_flux_record_host(double gamma, const double* u, double* f, double* g, long long ncells, int stride);

/* ------------------------------------------------------------------------
 * Batch engine: many independent solves (different grids, geometries, inflow
 * states and step counts) served by one device-resident loop. No counterpart in
 * the reference, which runs one solve per process.
 *
 * Requests queue on the host (FIFO). Each ibtfo_batch_progress call
 *   - replays a captured HIP graph of `chunk` batched steps for the admitted
 *     solves (one launch per kernel of the step sequence covers every solve),
 *   - meanwhile admits queued requests whose device memory fits the remaining
 *     HBM budget (an arena reserved at creation: no hipMalloc per request) and
 *     builds their geometry and initial state on a second stream, for as long
 *     as the chunk runs (at least one; up to 8 when no solve is active),
 *   - then retires the solves that reached their step count, downloading the
 *     state of those that asked for it and returning their memory.
 * All solves of one batch share the kernel selection of ibtfo_batch_config;
 * each request brings its own ibtfo_config (grid, extents, CFL, edge policy and
 * inflow state; max_iter = its number of steps; Psy_time ends it early).
 * ------------------------------------------------------------------------ */
typedef struct ibtfo_batch ibtfo_batch;

typedef struct ibtfo_batch_config {
    int scheme, riemann, limiter, boundary, time, precision, quirks; /* shared kernel selection */
    long long hbm_budget;  /* bytes of device memory for admitted solves */
    int max_slots;         /* solves resident at once */
    int max_step_blocks;   /* step-kernel grid capacity (strips of 60 columns x rows) */
    int max_ghost_blocks;  /* ghost-fill grid capacity (128 ghosts per block) */
    int max_ring_blocks;   /* edge-ring grid capacity (256 ring cells per block) */
    int rows;              /* strip height of every solve (0: 16) */
    int chunk;             /* steps per graph replay */
    int persistent;        /* 1: each replay is one persistent kernel that runs the chunk's steps of
                              every resident solve, its phases separated by grid barriers (the grid
                              sized to what the device holds resident); 0: one launch per phase */
} ibtfo_batch_config;

typedef struct ibtfo_request {
    ibtfo_config cfg;      /* the solve; slot fields must equal the batch's */
    int ls_kind;           /* level-set hook (ibtfo_levelset_kind) and its parameters */
    int ls_nprm;
    const double* ls_prm;
    const double* phi;     /* IBTFO_LS_FIELD: M*N values, else NULL */
    int init_kind;         /* initialiser (ibtfo_init_kind) and its parameters */
    int init_nprm;
    const double* init_prm;
    int keep_state;        /* 1: keep the final state for ibtfo_batch_result */
    double arrival;        /* arrival time on the batch clock (s); < 0: now */
} ibtfo_request;

typedef struct ibtfo_result {
    int status;            /* IBTFO_OK, or IBTFO_E_NONFINITE if a state went bad */
    int iter;              /* steps taken */
    double now_t, dt, mass;
    int M, N;              /* state dimensions (with ghosts) */
    double t_arrival, t_admit, t_done; /* batch clock (s) */
} ibtfo_result;

typedef struct ibtfo_batch_stats {
    int active, queued, finished;
    long long chunks, steps, cell_updates; /* interior cells x steps of finished and active solves */
    long long arena_used, arena_capacity;
    double busy_s;         /* time spent in progress calls */
} ibtfo_batch_stats;

/* Default batch settings: FP64 MUSCL / HLL / minmod / Regular, 8 GiB budget,
 * 256 slots, capacities for that budget, rows 16, chunk 50, one launch per phase. */
int ibtfo_batch_config_default(ibtfo_batch_config* bc);
/* Reserves the arena, the device tables and two non-blocking streams. */
int ibtfo_batch_create(const ibtfo_batch_config* bc, ibtfo_batch** out);
int ibtfo_batch_destroy(ibtfo_batch* b);
/* Tile autotuner of the batch engine: with the batch idle (nothing resident or
 * queued), makes the sample requests resident (as many as fit; each run without a
 * stop for the probe), times `probe_chunks` graph replays of a chunk for each
 * strip height (4, 8, 16, 32 rows) with one launch per phase and with the
 * persistent kernel, discards the sample, and sets the batch's rows and mode to
 * the pair with the most cell-updates/s (a shorter strip needs more step blocks
 * per solve, so fewer of the sample may fit the grid capacity). The best rows,
 * mode and milliseconds per batched step are returned (pointers may be NULL). */
int ibtfo_batch_autotune(ibtfo_batch* b, const ibtfo_request* sample, int n, int probe_chunks, int* best_rows,
                         int* best_persistent, double* best_ms_per_step);
/* Seconds since the batch was created (the clock of arrivals and results). */
double ibtfo_batch_clock(const ibtfo_batch* b);
/* Queues a request (parameters are copied); *ticket identifies it. Rejects
 * requests that could never fit (IBTFO_E_NOMEM) or disagree with the batch's
 * kernel selection (IBTFO_E_ARG). */
int ibtfo_batch_submit(ibtfo_batch* b, const ibtfo_request* r, long long* ticket);
/* One scheduling round (see above). Counts may be NULL. */
int ibtfo_batch_progress(ibtfo_batch* b, int* active, int* queued, int* finished);
/* Pops the ticket of a retired request: returns 1 and sets *ticket, or 0. */
int ibtfo_batch_poll(ibtfo_batch* b, long long* ticket);
/* Result of a retired request; U_OLD (M*N*4 doubles, AoS) only if it kept its state. */
int ibtfo_batch_result(ibtfo_batch* b, long long ticket, ibtfo_result* res, double* U_OLD);
/* Frees the host copy of a retired request's result. */
int ibtfo_batch_release(ibtfo_batch* b, long long ticket);
int ibtfo_batch_stats_get(ibtfo_batch* b, ibtfo_batch_stats* st);

/* Last error message of this thread (static storage). */
const char* ibtfo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* IBTFO_H_ */
