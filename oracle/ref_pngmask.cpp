// Harness for the reference's PNG mask level set (TEST INFRASTRUCTURE ONLY,
// linked into oracle/_ref/libref_pngmask.so by oracle/build_ref.py with the
// reference's Test/PNGGeo/Level_Set.cpp, Src/Utils/PngProcess.cpp,
// Src/Utils/ParamReader.cpp and Src/Geometry/Psy_coord.cpp, against the image's
// libpng). It exists so the oracle's mask level set and the library's PNG
// decoder can be checked bit for bit against the reference's own code.
//
// Test/PNGGeo/Level_Set.cpp:11-364 keeps the single-device signature (no
// ndevices/device) and reads "Demo.png" from the working directory; its
// per-pixel scratch (new double[W*H*6], :33) leaves the gradient slot of the
// border pixels unset and then scans every pixel's slot (:82-95), so stale heap
// contents can add edge points (and overrun the edge list). build_ref.py links
// this library with --wrap for operator new[] / delete[]: the reference's
// arrays come from calloc, so those slots read 0 (no edge), deterministically.
#include <unistd.h>

#include <cstdlib>
#include <cstring>

#include <png.h>

png_bytep* read_png_file(char* filename, int* width, int* height);  // Src/Utils/PngProcess.H:5
void Level_Set(char* filename, const int N_x, const int N_y, const int num_ghost_cell, double* XYCOORD);
void Psy_coord(const double lo_x, const double lo_y, const double Psy_L, const double Psy_H, const int N_x,
               const int N_y, const int num_ghost_cell, double* XYCOORD, const int ndevices, const int device);

extern "C" {

// operator new[] / delete[] of the reference objects linked into this library
void* __wrap__Znam(size_t n) {
    void* p = calloc(n ? n : 1, 1);
    if (!p) abort();
    return p;
}
void __wrap__ZdaPv(void* p) { free(p); }

// The reference reader's output (8-bit RGBA rows, top row first) into rgba
// (cap bytes). Returns 0, or -1 when it does not fit.
int ref_png_read(const char* path, unsigned char* rgba, long long cap, int* width, int* height) {
    char* p = strdup(path);
    png_bytep* rows = read_png_file(p, width, height);
    free(p);
    const long long rowbytes = 4ll * *width;
    int rc = 0;
    if (rowbytes * *height > cap) rc = -1;
    for (int y = 0; y < *height; y++) {
        if (rc == 0) memcpy(rgba + rowbytes * y, rows[y], rowbytes);
        free(rows[y]);
    }
    free(rows);
    return rc;
}

// Psy_coord, then the PNGGeo Level_Set over dir/Demo.png with the .inp file
// `inp` (lo/hi extents); XY gets the reference's XYCOORD (M*N*6 doubles).
int ref_mask_level_set(const char* dir, const char* inp, int N_x, int N_y, int ng, double lo_x, double lo_y,
                       double hi_x, double hi_y, double* XY) {
    const int M = N_x + 2 * ng, N = N_y + 2 * ng;
    memset(XY, 0, sizeof(double) * 6 * (size_t)M * N);
    Psy_coord(lo_x, lo_y, hi_x - lo_x, hi_y - lo_y, N_x, N_y, ng, XY, 1, 0);
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(dir) != 0) return -2;
    char* f = strdup(inp);
    Level_Set(f, N_x, N_y, ng, XY);
    free(f);
    return chdir(cwd) == 0 ? 0 : -3;
}

}  // extern "C"
