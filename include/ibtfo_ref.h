/*
 * ibtfo_ref.h -- the reference's own entry points over the MI355X library, for
 * drivers written against IBTFO (Test/<Case>/<Case>.cpp) that keep their call
 * sequence. Implemented in ibtfo_amd/csrc/ibtfo_refapi.cpp, exported by
 * libibtfo_amd.so.
 *
 * The four C functions replace the driver's OpenACC data directives
 * (Test/Sod/Sod.cpp:74-78 enter data, :88-92 update device, :100-104 update
 * self, :187-191 exit data): they create the solve from the .inp file and bind
 * it to the driver's host arrays, which the C++ functions below then find by
 * pointer (the OpenACC present table's role). Each returns an IBTFO_* status.
 */
#ifndef IBTFO_REF_H_
#define IBTFO_REF_H_

#ifdef __cplusplus
extern "C" {
#endif

/* acc enter data create(U_OLD, U_NEW, XYCOORD, SCHEME_IDX): one solve per call.
 * Arrays are the driver's host arrays in the reference's AoS layout. */
int ibtfo_ref_enter(const char* inp, double* U_OLD, double* U_NEW, double* XYCOORD, double* SCHEME_IDX);
/* acc update device(U_OLD, U_NEW) */
int ibtfo_ref_update_device(double* U_OLD);
/* acc update self(U_OLD) */
int ibtfo_ref_update_self(double* U_OLD);
/* acc exit data: destroys the solve bound to U_OLD */
int ibtfo_ref_exit(double* U_OLD);

#ifdef __cplusplus
}

/* The reference's free functions (C++ linkage, reference signatures). A failing
 * call prints the library's message and aborts: these signatures have no error
 * channel. ndevices must be 1 and device 0. */
void Psy_coord(const double lo_x, const double lo_y, const double Psy_L, const double Psy_H, const int N_x,
               const int N_y, const int num_ghost_cell, double* XYCOORD, const int ndevices,
               const int device);                                            /* Src/Geometry/Psy_coord.H:11-20 */
void Level_Set(char* filename, const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD,
               const int ndevices, const int device);                       /* Src/Geometry/Level_Set.H:12-18 */
void Scheme_Index(const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD, double* SCHEME_IDX,
                  const int ndevices, const int device);           /* Src/BoundaryCondition/Scheme_Index.H:13-19 */
void Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
              double* U_NEW, double* XYCOORD, double* SCHEME_IDX, const int ndevices,
              const int device);                                    /* Src/BoundaryCondition/Boundary.H:11-20 */
void Global_Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
                     double* U_NEW, const int ndevices, const int device); /* Src/BoundaryCondition/Boundary.H:22-29 */
void Advance(const double Psy_L, const double Psy_H, const int N_x, const int N_y, const int num_ghost_cell,
             const double gamma, double dt, double* U_OLD, double* U_TMP, double* U_NEW, double* XYCOORD,
             const int ndevices, const int device);    /* Src/Base/Euler/Advection/Advance.H:14-26 */
void TimeAdvance(const double Psy_L, const double Psy_H, const int N_x, const int N_y, const int num_ghost_cell,
                 const double gamma, double dt, double* U_OLD, double* U_TMP, double* U_NEW, double* XYCOORD,
                 double* SCHEME_IDX, const int ndevices,
                 const int device);                          /* Src/Base/TimeAdvance/TimeAdvance.H:13-26 */
/* The scheme's inner operators on a host scratch record U_TMP (M*N cells x 28
 * doubles; copied to the GPU, computed there and copied back): */
void Riemann_solverX(const double Psy_L, const double Psy_H, const int N_x, const int N_y,
                     const int num_ghost_cell, const double gamma, double* U_OLD, double* U_TMP,
                     const int ndevices, const int device); /* Src/Base/Euler/Riemann/Riemann_solver.H:22-31 */
void Riemann_solverY(const double Psy_L, const double Psy_H, const int N_x, const int N_y,
                     const int num_ghost_cell, const double gamma, double* U_OLD, double* U_TMP,
                     const int ndevices, const int device); /* Src/Base/Euler/Riemann/Riemann_solver.H:33-42 */
void Conserve2Flux(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U, double* F,
                   double* G, const int ndevices, const int device);  /* Src/Base/Common/Conserve2Flux.H:11-19 */
void Conserve2FluxX(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U,
                    double* F, const int ndevices, const int device); /* Src/Base/Common/Conserve2Flux.H:21-28 */
void Conserve2FluxY(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U,
                    double* G, const int ndevices, const int device); /* Src/Base/Common/Conserve2Flux.H:30-37 */
void ComputeDt(const double Psy_L, const double Psy_H, const int N_x, const int N_y, const int num_ghost_cell,
               const double gamma, const double CFL_number, double* U_OLD, double* XYCOORD, double* TMP_DT,
               const int ndevices, const int device);                  /* Src/Base/Common/ComputeDt.H:12-23 */
void CaculateMass(const double Psy_L, const double Psy_H, const int N_x, const int N_y, const int num_ghost_cell,
                  double* U_OLD, double* XYCOORD, double* MASS_DEVICE, const int ndevices,
                  const int device);                                     /* Src/Utils/CaculateMass.H:14-23 */
void WriteData(const double lo_x, const double lo_y, const double Psy_L, const double Psy_H, const int N_x,
               const int N_y, const int num_ghost_cell, int iter, double now_t, const double gamma, double* U_OLD,
               double* XYCOORD, const int ndevices, const int device);    /* Src/Utils/WriteData.H:14-27 */
#endif

#endif /* IBTFO_REF_H_ */
