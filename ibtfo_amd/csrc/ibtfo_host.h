// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <cstddef>
#include <map>
#include <string>

struct ibtfo_config;  // include/ibtfo.h

namespace ibtfo {

void set_error(const char* fmt, ...);
std::string first_number(const std::string& s);
bool inp_numbers(const char* path, std::map<std::string, double>& out);
bool inp_strings(const char* path, std::map<std::string, std::string>& out);

// Initialize hooks evaluated on the host (AoS output, reference cell order);
// x/y/type per cell come from the geometry.
int host_initialize(int kind, const double* prm, int nprm, int nx, int ny, int ng, double gam, const double* x,
                    const double* y, const unsigned char* type, double* U);

// WriteData's files for host arrays in the reference's AoS layout (ibtfo_vtk.cpp)
int write_vtk(const ::ibtfo_config* c, int iter, double now_t, const double* U, const double* XY,
              const char* dir);

// First-fit sub-allocator over one device allocation: the batch engine's HBM
// budget. Admission and retirement of solves allocate from it without any
// hipMalloc / hipFree (which would synchronise the device). Not thread-safe.
class Arena {
  public:
    int init(size_t bytes);  // one hipMalloc of `bytes`; IBTFO_* status
    void release();
    void* alloc(size_t bytes);  // 256-B aligned; nullptr when nothing fits
    void free(void* p);
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t largest_free() const;

  private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> live_;  // offset -> size
};

}  // namespace ibtfo
