// Signature adapter for the reference's Test/PNGGeo hooks (TEST INFRASTRUCTURE
// ONLY, linked into oracle/_ref/libref_pngshock_*.so by oracle/build_ref.py).
// Test/PNGGeo/Initialize.cpp:3-13 and Test/PNGGeo/Global_Boudary.cpp:7-12 keep
// the single-device signatures (no ndevices/device arguments); this file
// supplies the current ones and forwards to the reference's own definitions.
void Initialize(char* filename, const double Psy_L, const double Psy_H, const int N_x, const int N_y,
                const int num_ghost_cell, const double gamma, double* U_OLD, double* U_NEW, double* XYCOORD);
void Global_Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
                     double* U_NEW);

void Initialize(char* filename, const double Psy_L, const double Psy_H, const int N_x, const int N_y,
                const int num_ghost_cell, const double gamma, double* U_OLD, double* U_NEW, double* XYCOORD,
                const int, const int) {
    Initialize(filename, Psy_L, Psy_H, N_x, N_y, num_ghost_cell, gamma, U_OLD, U_NEW, XYCOORD);
}

void Global_Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
                     double* U_NEW, const int, const int) {
    Global_Boundary(N_x, N_y, num_ghost_cell, gamma, U_OLD, U_NEW);
}
