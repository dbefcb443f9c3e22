// Signature adapter for the reference's SimpleBound directory (TEST
// INFRASTRUCTURE ONLY, linked into oracle/_ref/libref_*_SimpleBound.so by
// oracle/build_ref.py). SimpleBound/Scheme_Index.cpp:5-9 and
// SimpleBound/LS_SimpleBound.cpp:6-13 still define the single-device signatures
// (no ndevices/device arguments), so against the current headers they compile as
// separate overloads and a driver calling the current ones does not link. This
// file supplies the current signatures and forwards to the reference's own
// single-device definitions, unchanged.
void Scheme_Index(const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD, double* SCHEME_IDX);
void Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
              double* U_NEW, double* XYCOORD, double* SCHEME_IDX);

void Scheme_Index(const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD, double* SCHEME_IDX,
                  const int, const int) {
    Scheme_Index(N_x, N_y, num_ghost_cell, XYCOORD, SCHEME_IDX);
}

void Boundary(const int N_x, const int N_y, const int num_ghost_cell, const double gamma, double* U_OLD,
              double* U_NEW, double* XYCOORD, double* SCHEME_IDX, const int, const int) {
    Boundary(N_x, N_y, num_ghost_cell, gamma, U_OLD, U_NEW, XYCOORD, SCHEME_IDX);
}
