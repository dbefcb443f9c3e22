// Geometry set-up kernels (ibtfo_geom.hip).
#pragma once
#include "ibtfo_internal.h"

namespace ibtfo {

// shape record: {type (0 circle, 1 convex polygon), cx, cy, r, nverts, x0, y0, ... x7, y7}
constexpr int SHAPE_STRIDE = 21;

struct GeomArgs {
    Layout L;
    double Li, Hi, dx;      // first cell centre (with ghosts) and spacing, Psy_coord.cpp:21-26
    int kind;               // level-set hook
    double p[8];            // hook constants (evaluated on the host)
    const double* phi_in;   // kind 3: per-cell phi in reference cell order
    const double* shapes;   // kind 5: nshapes records of SHAPE_STRIDE doubles
    int nshapes;
    double *phi, *nx, *ny;  // planes
    uint8_t* type;
    int bc;
    int unit_normals;       // corrected mode: normals scaled to unit length
    // cut cells (bc = IDW fill with active donors = types 0 and 3)
    int cut;
    double cdx, cdy;        // cell sizes (Psy_L / N_x, Psy_H / N_y)
    double* dcut;           // dense scratch: M x N records of 8 doubles
    CutCell* cuts;          // compact records of the type-3 cells
    int* cut_index;         // plane: record of each type-3 cell
    unsigned int* ccount;
    int ccap;
    MergeGroup* groups;
    unsigned int* gcount;
    int gcap;
    Ghost* ghosts;
    unsigned int* count;
    int cap;
};

// Mask level set (Test/PNGGeo/Level_Set.cpp:31-149): phi per cell from an RGBA image
struct MaskArgs {
    const unsigned char* rgba;  // W x H x 4, top row first
    int W, H;
    double* v;                  // W x H signs, bottom row first
    unsigned char* edge;        // W x H edge flags, bottom row first
    int M, N;                   // cells with ghosts
    double Li, Hi, dxc;         // cell centres (Psy_coord)
    double exo, eyo, dxp, dyp;  // edge pixel x: exo + dxp * x, y: eyo + dyp * y
    double* g1;                 // N x W (cell row major): squared y distance to the column's nearest edge
    int* r1;                    // N x W: that edge's pixel row (-1: none)
    int* r2;                    // N x W: the column's other candidate row (-1: none)
    int nseg;                   // 32-row segments per pixel column
    int *sfirst, *slast;        // nseg x W: first / last edge row of each segment (-1: none)
    int *before, *after;        // nseg x W: last edge row below / first above each segment (-1: none)
    int* vbuf;                  // N x (W+1) envelope scratch (masks too wide for the LDS envelope)
    double* zbuf;               // N x (W+1)
    double* phi;                // M x N, reference cell order
};
constexpr int MASK_SEG = 32;
// LDS envelope of one cell row: z (W+1 doubles), g (W doubles), v (W+1 uint16)
inline size_t mask_lds_bytes(int W) { return (size_t)(W + 1) * 8 + (size_t)W * 8 + (size_t)(W + 1) * 2 + 16; }
constexpr size_t MASK_LDS_MAX = 160 * 1024;

hipError_t geom_run(GeomArgs& g, hipStream_t s);
hipError_t mask_run(MaskArgs& m, hipStream_t s);  // row envelope in LDS when m.vbuf is null
const void* mask_rows_kernel();                     // that kernel (its LDS attribute is set by the caller)
// cut-cell geometry on top of geom_run's phi and normals: records, types, merge
// targets, ghosts (dense scratch g.dcut), then the compaction passes (counting
// when ccap / gcap are 0)
hipError_t cut_run(GeomArgs& g, hipStream_t s);
hipError_t cut_compact(GeomArgs& g, hipStream_t s);
hipError_t geom_ghosts(GeomArgs& g, hipStream_t s);

}  // namespace ibtfo
