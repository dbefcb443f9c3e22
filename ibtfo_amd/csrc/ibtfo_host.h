// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <map>
#include <string>

namespace ibtfo {

void set_error(const char* fmt, ...);
std::string first_number(const std::string& s);
bool inp_numbers(const char* path, std::map<std::string, double>& out);
bool inp_strings(const char* path, std::map<std::string, std::string>& out);

// Initialize hooks evaluated on the host (AoS output, reference cell order);
// x/y/type per cell come from the geometry.
int host_initialize(int kind, const double* prm, int nprm, int nx, int ny, int ng, double gam, const double* x,
                    const double* y, const unsigned char* type, double* U);

}  // namespace ibtfo
