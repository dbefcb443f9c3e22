// Output in the reference's on-disk format (Src/Utils/WriteData.cpp:5-247): one
// legacy-ASCII VTK STRUCTURED_POINTS file per output, dat_<device>_<iter>.vtk,
// with the interior cells' density, x-vel, y-vel, pressure, Energy, Mach_Number
// and velocity as "%f" lines, and the index file result.visit, appended to
// ("!NBLOCKS <ndevices>" before the first entry of iteration 0). One device
// here, so the file covers the whole domain (device 0, ndevices 1).
//
// The text is formatted on the host (the file is host output either way), rows
// split over threads into per-thread buffers written in order, so the bytes are
// those of the reference's sequential fprintf loops.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ibtfo.h"
#include "ibtfo_host.h"

namespace ibtfo {

namespace {

// one field of the interior, rows j0 <= j < j1, appended to out
template <typename F> void rows_text(std::string& out, int M, int ng, int nx, int j0, int j1, F value) {
    char buf[512];
    for (int j = j0; j < j1; j++)
        for (int i = ng; i < nx + ng; i++) {
            int n = snprintf(buf, sizeof buf, "%f  \n", value((size_t)M * j + i));
            out.append(buf, n > 0 ? std::min(n, (int)sizeof buf - 1) : 0);
        }
}

}  // namespace

int write_vtk(const ibtfo_config* c, int iter, double now_t, const double* U, const double* XY, const char* dir) {
    const int nx = c->N_x, ny = c->N_y, ng = c->num_ghost_cell, M = nx + 2 * ng;
    const double lo_x = c->lo_x, lo_y = c->lo_y, L = c->hi_x - c->lo_x, H = c->hi_y - c->lo_y, gam = c->gamma;
    const std::string base = (dir && dir[0]) ? std::string(dir) + "/" : std::string();

    FILE* idx = fopen((base + "result.visit").c_str(), "a");
    if (!idx) { set_error("cannot open %sresult.visit", base.c_str()); return IBTFO_E_IO; }
    if (iter == 0) fprintf(idx, "!NBLOCKS %d\n", 1);
    fprintf(idx, "dat_%d_%d.vtk\n", 0, iter);
    fclose(idx);

    const std::string name = base + "dat_0_" + std::to_string(iter) + ".vtk";
    FILE* out = fopen(name.c_str(), "w");
    if (!out) { set_error("cannot open %s", name.c_str()); return IBTFO_E_IO; }
    const int j0 = ng, j1 = ny + ng;  // one device: every interior row
    fprintf(out, "# vtk DataFile Version 2.0 \n");
    fprintf(out, "Volume example \n");
    fprintf(out, "ASCII \n");
    fprintf(out, "DATASET STRUCTURED_POINTS \n");
    fprintf(out, "DIMENSIONS %d %d %d  \n", nx, j1 - j0, 1);
    fprintf(out, "ASPECT_RATIO 1 1 1 \n");
    fprintf(out, "ORIGIN %f %f %f \n", lo_x + L / nx / 2, lo_y + H / ny / 2 + (j0 - ng) * H / ny, 0.0);
    fprintf(out, "SPACING %f %f %f \n", L / nx, H / ny, H / ny);
    fprintf(out, "FIELD FieldData 1  \n");
    fprintf(out, "TIME 1 1 double  \n");
    fprintf(out, "%f  \n", now_t);
    fprintf(out, "POINT_DATA  %d\n", nx * (j1 - j0));

    auto fluid = [&](size_t q) { return XY[q * 6 + 5] < 0.5; };
    auto prim = [&](size_t q, double& rho, double& u, double& v, double& p) {
        rho = U[q * 4];
        u = U[q * 4 + 1] / U[q * 4];
        v = U[q * 4 + 2] / U[q * 4];
        p = (gam - 1) * (U[q * 4 + 3] - 0.5 * rho * (u * u + v * v));
    };
    // the seven fields, each with its header lines as the reference prints them
    auto density = [&](size_t q) { return fluid(q) ? (U[q * 4] > 1e-6 ? U[q * 4] : 1e-6) : -1.0; };
    auto xvel = [&](size_t q) { return fluid(q) ? U[q * 4 + 1] / U[q * 4] : -100.0; };
    auto yvel = [&](size_t q) { return fluid(q) ? U[q * 4 + 2] / U[q * 4] : -100.0; };
    auto pres = [&](size_t q) {
        if (!fluid(q)) return -1.0;
        double rho, u, v, p;
        prim(q, rho, u, v, p);
        return p;
    };
    auto energy = [&](size_t q) { return fluid(q) ? U[q * 4 + 3] : -1.0; };
    auto mach = [&](size_t q) {
        if (!fluid(q)) return -1.0;
        double rho, u, v, p;
        prim(q, rho, u, v, p);
        double a = pow((gam * p / rho), 0.5);
        return sqrt(u * u + v * v) / a;
    };
    auto speed = [&](size_t q) {
        if (!fluid(q)) return -1.0;
        double u = U[q * 4 + 1] / U[q * 4], v = U[q * 4 + 2] / U[q * 4];
        return sqrt(u * u + v * v);
    };

    const int rows = j1 - j0;
    const int nthreads = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    auto emit = [&](const char* header, auto fn) -> bool {
        fputs(header, out);
        const int nt = std::min(nthreads, std::max(1, rows / 8));
        std::vector<std::string> part(nt);
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t]() {
                const int a = j0 + (int)((long long)rows * t / nt), b = j0 + (int)((long long)rows * (t + 1) / nt);
                part[t].reserve((size_t)(b - a) * nx * 12);
                rows_text(part[t], M, ng, nx, a, b, fn);
            });
        for (auto& x : th) x.join();
        for (auto& s : part)
            if (fwrite(s.data(), 1, s.size(), out) != s.size()) return false;
        return true;
    };
    bool ok = emit("SCALARS density float  \nLOOKUP_TABLE default  \n", density) &&
              emit("SCALARS x-vel float \nLOOKUP_TABLE default  \n", xvel) &&
              emit("SCALARS y-vel float  \nLOOKUP_TABLE default  \n", yvel) &&
              emit("SCALARS pressure float  \nLOOKUP_TABLE default  \n", pres) &&
              emit("SCALARS Energy float  \nLOOKUP_TABLE default  \n", energy) &&
              emit("SCALARS Mach_Number float  \nLOOKUP_TABLE default  \n", mach) &&
              emit("SCALARS velocity float  \nLOOKUP_TABLE default  \n", speed);
    if (fclose(out) != 0 || !ok) { set_error("write error on %s", name.c_str()); return IBTFO_E_IO; }
    return IBTFO_OK;
}

}  // namespace ibtfo

extern "C" int ibtfo_vtk_write_host(const ibtfo_config* cfg, int iter, double now_t, const double* U_OLD,
                                    const double* XYCOORD, const char* dir) {
    if (!cfg || !U_OLD || !XYCOORD) { ibtfo::set_error("ibtfo_vtk_write_host: null argument"); return IBTFO_E_ARG; }
    return ibtfo::write_vtk(cfg, iter, now_t, U_OLD, XYCOORD, dir);
}
