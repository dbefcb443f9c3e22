// The solve context behind the C ABI's opaque ibtfo_ctx, shared by the C-ABI
// translation unit and the batch engine.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/ibtfo.h"
#include "ibtfo_host.h"
#include "ibtfo_internal.h"

struct ibtfo_ctx {
    // one solve: its configuration, device memory, launch arguments and stream
    ibtfo_config cfg;
    ibtfo::Layout L;
    int device;
    hipStream_t stream;
    hipStream_t own_stream;
    double* state[2];
    double* unew0;  // SimpleBound: the uploaded U_NEW (its non-fluid cells stay as uploaded)
    uint8_t* type;
    double *phi, *gnx, *gny;
    ibtfo::Ghost* ghosts;
    int nghost;
    int nlevels;  // ghost-fill passes per Boundary (SimpleBound dependency levels; else 1)
    double* gvals;
    unsigned int* dcount;
    ibtfo::Ctl* ctl;
    double* mass_part;  // PART_STRIDE doubles per step-kernel block
    int mass_cap;       // blocks
    double* fold;       // FOLD_MAX records of the two-level finish
    int rows_override;  // step-kernel rows per strip (0 = default shape)
    int rowpar;         // 1: the row-parallel step kernel (rows + 1 waves per block)
    double* corner;
    double Li, Hi, dxc;  // cell-centre origin and spacing (Psy_coord)
    int have_geometry, have_state, prepared;
    ibtfo::StepArgs sa;
    ibtfo::FillArgs fa;
    ibtfo::EdgeArgs ea;
    ibtfo::ReduceArgs ra;
    ibtfo::FinishArgs fin;
    ibtfo::MergeArgs ma;
    ibtfo::CutCell* cuts;     // cut-cell records, index plane and merge groups
    int* cut_index;
    ibtfo::MergeGroup* groups;
    int ncut, ngroups;
    uint8_t* stype;       // the cut-cell step kernel's type plane: type | MERGED_BIT for group cells
    double* merge_part;   // PART_STRIDE doubles per merge block
    double* df_part;      // deferred finish: 2 x DF_MAX_REC partial records (single solves only)
    hipGraphExec_t graph;
    int graph_chunk;
    std::vector<double> corner_host;
    ibtfo::Arena* arena;  // device memory from a batch engine's arena (else hipMalloc)
};

namespace ibtfo {
// Creates a context on `stream` (own non-blocking stream when null) with device
// memory from `arena` (hipMalloc when null); rows = strip height (0 = default).
int ctx_create(const ibtfo_config* cfg, Arena* arena, hipStream_t stream, int rows, ibtfo_ctx** out);
void free_ctx(ibtfo_ctx* x);   // releases device memory and the context (no synchronisation)
void build_args(ibtfo_ctx* x); // (re)derives the kernel arguments from the context
}  // namespace ibtfo
