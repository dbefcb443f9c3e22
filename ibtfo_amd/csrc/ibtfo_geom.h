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
    Ghost* ghosts;
    unsigned int* count;
    int cap;
};

hipError_t geom_run(GeomArgs& g, hipStream_t s);
hipError_t geom_ghosts(GeomArgs& g, hipStream_t s);

}  // namespace ibtfo
