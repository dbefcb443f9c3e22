// C entry point for the reference's WriteData (Src/Utils/WriteData.cpp:5-247),
// compiled with it from the reference sources into oracle/_ref/libref_writedata.so
// by oracle/build_ref.py. TEST INFRASTRUCTURE ONLY: the byte-level check of
// the library's VTK output runs both writers on the same state.
void WriteData(const double lo_x, const double lo_y, const double Psy_L, const double Psy_H, const int N_x,
               const int N_y, const int num_ghost_cell, int iter, double now_t, const double gamma, double* U_OLD,
               double* XYCOORD, const int ndevices, const int device);

extern "C" void ref_write_data(double lo_x, double lo_y, double Psy_L, double Psy_H, int N_x, int N_y, int ng,
                               int iter, double now_t, double gamma, double* U_OLD, double* XYCOORD) {
    WriteData(lo_x, lo_y, Psy_L, Psy_H, N_x, N_y, ng, iter, now_t, gamma, U_OLD, XYCOORD, 1, 0);
}
