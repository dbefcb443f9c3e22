// ibtfo_run: first-party driver with the reference drivers' call order
// (Test/Sod/Sod.cpp:94-181, Test/InclineSod/InclineSod.cpp:73-165) on the
// MI355X library: read the .inp, geometry, Initialize, [Boundary], ComputeDt,
// then the step loop, printing the reference's per-step line.
//
//   ibtfo_run case.inp [--chunk K] [--quiet] [--vtk DIR]
//
// With --vtk, WriteData's files go to DIR at the reference driver's cadence
// (Test/Sod/Sod.cpp:135-150, :186): every plot_per steps (plot_int = 0), or
// whenever now_t passes the next of plot_int equal time marks, and at the end.
//
// Problem hooks come from .inp keys (the reference picks them by which test
// directory is compiled): Case = sod | sodx | incline | vortex | uniform | shu |
// pnggeo (Test/PNGGeo: the mask PNG named by mask_file, default Demo.png, and
// the Mach 1.3 shock initialiser; its edge hook is Edge = pnggeo).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ibtfo.h"

static double num(const char* inp, const char* key, double def) {
    double v = def;
    int r = ibtfo_inp_number(inp, key, &v);
    if (r == 0) printf("missing param \"%s\",use default value %g\n", key, def);
    return v;
}

static std::string str(const char* inp, const char* key, const char* def) {
    char buf[256];
    int n = ibtfo_inp_string(inp, key, buf, sizeof buf);
    if (n <= 0) return def;
    std::string s(buf);
    size_t h = s.find('#');
    if (h != std::string::npos) s = s.substr(0, h);
    return s;
}

#define CK(x)                                                                       \
    do {                                                                            \
        int rc_ = (x);                                                              \
        if (rc_) {                                                                  \
            fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, ibtfo_last_error());   \
            return 1;                                                               \
        }                                                                           \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s case.inp [--chunk K] [--quiet] [--vtk DIR]\n", argv[0]);
        return 2;
    }
    const char* inp = argv[1];
    int chunk = 64, quiet = 0;
    const char* vtk = nullptr;
    for (int a = 2; a < argc; a++) {
        if (!strcmp(argv[a], "--chunk") && a + 1 < argc) chunk = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--quiet")) quiet = 1;
        else if (!strcmp(argv[a], "--vtk") && a + 1 < argc) vtk = argv[++a];
    }
    if (chunk < 1) chunk = 1;
    ibtfo_config cfg;
    CK(ibtfo_config_from_inp(inp, &cfg));
    printf(" ==== parameters are read ====\n");
    ibtfo_ctx* ctx = nullptr;
    CK(ibtfo_create(&cfg, &ctx));
    printf(" ====  memory allocation complete ====\n");

    std::string cs = str(inp, "Case", cfg.boundary == IBTFO_REGULAR ? "sod" : "incline");
    int ls = IBTFO_LS_NONE, init = IBTFO_INIT_SOD_ROWS;
    std::vector<double> lp, ip;
    std::vector<double> L = {num(inp, "rho_L", 1), num(inp, "u_L", 0), num(inp, "v_L", 0), num(inp, "p_L", 1)};
    std::vector<double> R = {num(inp, "rho_R", 0.125), num(inp, "u_R", 0), num(inp, "v_R", 0), num(inp, "p_R", 0.1)};
    if (cs == "sod") {
        ip = L; ip.insert(ip.end(), R.begin(), R.end());
    } else if (cs == "sodx") {
        init = IBTFO_INIT_SOD_X;
        ip = {num(inp, "x0", 0.5)}; ip.insert(ip.end(), L.begin(), L.end()); ip.insert(ip.end(), R.begin(), R.end());
    } else if (cs == "incline") {
        ls = IBTFO_LS_CHANNEL; init = IBTFO_INIT_INCLINE;
        lp = {num(inp, "angle", 30), num(inp, "AxisD", 0.4)};
        ip = {lp[0]}; ip.insert(ip.end(), L.begin(), L.end()); ip.insert(ip.end(), R.begin(), R.end());
    } else if (cs == "vortex") {
        ls = IBTFO_LS_ANNULUS; init = IBTFO_INIT_VORTEX;
        lp = {num(inp, "f0", 0), num(inp, "f1", 1)};
        ip = {num(inp, "fa", 0), num(inp, "fw", 0)};
    } else if (cs == "uniform") {
        init = IBTFO_INIT_UNIFORM;
        ip = {num(inp, "rho", 1), num(inp, "u", 0), num(inp, "v", 0), num(inp, "p", 1)};
    } else if (cs == "pnggeo") {
        init = IBTFO_INIT_SHOCK;
        ip = {0.1, 1.3};
        ip.insert(ip.end(), L.begin(), L.end());
    } else if (cs == "shu") {
        init = IBTFO_INIT_SHU_VORTEX;
        ip = {num(inp, "x0", 5), num(inp, "y0", 5), num(inp, "beta", 5), 1.0, num(inp, "u", 1), num(inp, "v", 1), 1.0};
    } else {
        fprintf(stderr, "unknown Case '%s'\n", cs.c_str());
        return 2;
    }
    if (cs == "pnggeo") {
        std::string png = str(inp, "mask_file", "Demo.png");
        CK(ibtfo_geometry_png(ctx, png.c_str()));
    } else {
        CK(ibtfo_geometry(ctx, ls, lp.data(), (int)lp.size(), nullptr));
    }
    printf(" ====  Geometry Initialize complete ====\n");
    CK(ibtfo_initialize(ctx, init, ip.data(), (int)ip.size()));
    CK(ibtfo_prepare(ctx, cfg.boundary != IBTFO_REGULAR));
    double mass0 = 0;
    CK(ibtfo_mass(ctx, &mass0));
    printf(" ====  Initial Mass is %g ====\n", mass0);

    int iter = 0;
    double now_t = 0, dt = 0;
    const int plot_per = vtk ? std::max(1, (int)num(inp, "plot_per", 100)) : 0;
    const int plot_int = vtk ? (int)num(inp, "plot_int", 0) : 0;
    int output_int = 0;
    std::vector<double> ldt(chunk), lt(chunk), lm(chunk);
    while (true) {
        int before = iter;
        int n_try = chunk;
        if (vtk) {
            if (plot_int == 0) {
                if (iter % plot_per == 0) {
                    printf(" ==== Writing Data Wait ====\n");
                    CK(ibtfo_write_vtk(ctx, vtk, iter, now_t));
                }
                n_try = std::min(chunk, plot_per - iter % plot_per);
            } else {
                if (now_t > output_int * cfg.Psy_time / plot_int || iter == 0) {
                    printf(" ==== Writing Data Wait ====\n");
                    CK(ibtfo_write_vtk(ctx, vtk, iter, now_t));
                    output_int = output_int + 1 * (iter > 0);
                }
                n_try = 1;  // the next mark depends on each step's dt
            }
        }
        CK(ibtfo_step_graph(ctx, n_try, chunk));
        int rc = ibtfo_status(ctx, &iter, &now_t, &dt);
        if (rc && rc != IBTFO_E_NONFINITE) CK(rc);
        int n = iter - before;
        if (n > 0 && !quiet) {
            CK(ibtfo_log(ctx, n, ldt.data(), lt.data(), lm.data()));
            for (int q = 0; q < n; q++)
                printf("iter : %d.   dt :%g.   time :%g.  Mass loss :%g\n", before + q + 1, ldt[q], lt[q],
                       std::fabs(lm[q] - mass0));
        }
        if (n < n_try) break;  // the device stopped the loop
    }
    if (vtk) CK(ibtfo_write_vtk(ctx, vtk, iter, now_t));
    printf(" ==== done: %d steps, t = %.9g ====\n", iter, now_t);
    CK(ibtfo_destroy(ctx));
    return 0;
}
