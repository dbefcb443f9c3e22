// .inp input files, with the reference's grammar (Src/Utils/ParamReader.cpp:14-68,
// Src/Utils/ParamReader.H:47-55):
//  - a line whose first character is '#' is skipped; every ' ' is removed (tabs
//    and '\r' are kept);
//  - key = text before the first '='; value = text after it (a line without '='
//    is keyed by the whole line and valued by all but its last character);
//  - numbers(): the first match of [+-]?(\d*\.\d+|\d+) in the value, through
//    atof ("" -> 0), so "1e-3" reads as 1 and inline "# comments" are ignored;
//  - strings(): the raw value, except values made only of digits;
//  - later duplicates win.
// Slot names (Scheme, Riemann, Limiter, Boundary, Time) are read through
// strings() and cut at the first '#', since the reference selects them in its
// Makefile, where '#' starts a comment.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/ibtfo.h"
#include "ibtfo_host.h"

namespace ibtfo {

static bool read_lines(const char* path, std::vector<std::string>& out) {
    std::ifstream in(path);
    if (!in.is_open()) return false;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::string t;
        t.reserve(line.size());
        for (char ch : line)
            if (ch != ' ') t.push_back(ch);
        out.push_back(t);
    }
    return true;
}

static void split_kv(const std::string& t, std::string& key, std::string& val) {
    size_t e = t.find('=');
    if (e == std::string::npos) {
        key = t;
        val = t.empty() ? std::string() : t.substr(0, t.size() - 1);
    } else {
        key = t.substr(0, e);
        val = t.substr(e + 1);
    }
}

static bool is_digit(char c) { return c >= '0' && c <= '9'; }

// leftmost match of [+-]?(\d*\.\d+|\d+) (first alternative preferred, greedy)
std::string first_number(const std::string& s) {
    const size_t n = s.size();
    for (size_t p = 0; p < n; p++) {
        const bool sign = (s[p] == '+' || s[p] == '-');
        for (int take = sign ? 1 : 0; take >= 0; take--) {
            size_t q = p + take, d = q;
            while (d < n && is_digit(s[d])) d++;
            if (d < n && s[d] == '.' && d + 1 < n && is_digit(s[d + 1])) {
                size_t e = d + 1;
                while (e < n && is_digit(s[e])) e++;
                return s.substr(p, e - p);
            }
            if (d > q) return s.substr(p, d - p);
        }
    }
    return std::string();
}

bool inp_numbers(const char* path, std::map<std::string, double>& out) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return false;
    for (const std::string& t : lines) {
        std::string k, v;
        split_kv(t, k, v);
        out[k] = atof(first_number(v).c_str());
    }
    return true;
}

bool inp_strings(const char* path, std::map<std::string, std::string>& out) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return false;
    for (const std::string& t : lines) {
        std::string k, v;
        split_kv(t, k, v);
        bool digits = !v.empty();
        for (char ch : v)
            if (!isdigit((unsigned char)ch)) { digits = false; break; }
        if (digits) continue;
        out[k] = v;
    }
    return true;
}

static int lookup(const std::map<std::string, std::string>& m, const char* key, const char* const* names,
                  const int* values, int n, int* dst) {
    auto it = m.find(key);
    if (it == m.end()) return 0;
    std::string v = it->second;
    size_t h = v.find('#');
    if (h != std::string::npos) v = v.substr(0, h);
    while (!v.empty() && (v.back() == '\r' || v.back() == '\t')) v.pop_back();
    for (int q = 0; q < n; q++)
        if (v == names[q]) { *dst = values[q]; return 0; }
    set_error("unknown %s '%s' in input file", key, v.c_str());
    return IBTFO_E_ARG;
}

}  // namespace ibtfo

using namespace ibtfo;

extern "C" int ibtfo_config_default(ibtfo_config* c) {
    if (!c) return IBTFO_E_ARG;
    memset(c, 0, sizeof *c);
    // Params::get defaults of the reference drivers, Test/Sod/Sod.cpp:34-48
    c->max_iter = 100000;
    c->lo_x = 0; c->lo_y = 0; c->hi_x = 1; c->hi_y = 1;
    c->N_x = 500; c->N_y = 200;
    c->Psy_time = 1.0;
    c->CFL_number = 0.3;
    c->num_ghost_cell = 2;
    c->gamma = 1.4;
    // Makefile defaults, Test/Sod/Makefile:1-5
    c->scheme = IBTFO_MUSCL;
    c->riemann = IBTFO_HLL;
    c->limiter = IBTFO_MINMOD;
    c->boundary = IBTFO_REGULAR;
    c->time = IBTFO_SINGLESTEP;
    c->edge = IBTFO_EDGE_REFERENCE;
    c->precision = IBTFO_FP64;
    c->quirks = 1;
    c->edge_state[0] = 1.0; c->edge_state[3] = 1.0;
    return IBTFO_OK;
}

extern "C" int ibtfo_config_from_inp(const char* path, ibtfo_config* c) {
    if (!path || !c) return IBTFO_E_ARG;
    ibtfo_config_default(c);
    std::map<std::string, double> num;
    std::map<std::string, std::string> str;
    if (!inp_numbers(path, num) || !inp_strings(path, str)) {
        set_error("cannot read %s", path);
        return IBTFO_E_IO;
    }
    auto get = [&](const char* k, double d) { auto it = num.find(k); return it == num.end() ? d : it->second; };
    c->max_iter = int(get("max_iter", c->max_iter));
    c->lo_x = get("lo_x", c->lo_x); c->lo_y = get("lo_y", c->lo_y);
    c->hi_x = get("hi_x", c->hi_x); c->hi_y = get("hi_y", c->hi_y);
    c->N_x = int(get("N_x", c->N_x)); c->N_y = int(get("N_y", c->N_y));
    c->Psy_time = get("Psy_time", c->Psy_time);
    c->CFL_number = get("CFL_number", c->CFL_number);
    c->num_ghost_cell = int(get("num_ghost_cell", c->num_ghost_cell));
    c->gamma = get("gamma", c->gamma);
    static const char* sn[] = {"Lax", "MUSCL", "MHM", "Godunov"};
    static const int sv[] = {IBTFO_LAX, IBTFO_MUSCL, IBTFO_MHM, IBTFO_GODUNOV};
    static const char* rn[] = {"HLL", "HLLC", "PVRS", "StegerWarming"};
    static const int rv[] = {IBTFO_HLL, IBTFO_HLLC, IBTFO_PVRS, IBTFO_STEGER_WARMING};
    static const char* ln[] = {"minmod", "VanLeer", "VanAlbada"};
    static const int lv[] = {IBTFO_MINMOD, IBTFO_VANLEER, IBTFO_VANALBADA};
    // OIDWBaseGFM is what Test/InclineSod/Makefile:4 and Test/Vortex/Makefile:4 name; no such
    // directory exists and IDWBaseGFM is the evident intent.
    static const char* bn[] = {"Regular", "IDWBaseGFM", "OIDWBaseGFM", "LinearBaseGFM", "RiemannBaseGFM", "BCGFM", "HBCGFM",
                               "CutCell", "SimpleBound"};
    static const int bv[] = {IBTFO_REGULAR, IBTFO_IDW_GFM, IBTFO_IDW_GFM, IBTFO_LINEAR_GFM, IBTFO_RIEMANN_GFM,
                             IBTFO_BCGFM, IBTFO_HBCGFM, IBTFO_CUTCELL, IBTFO_SIMPLE_BOUND};
    static const char* tn[] = {"SingleStep", "RK2"};
    static const int tv[] = {IBTFO_SINGLESTEP, IBTFO_RK2};
    static const char* en[] = {"reference", "periodic", "inflow", "pnggeo"};
    static const int ev[] = {IBTFO_EDGE_REFERENCE, IBTFO_EDGE_PERIODIC, IBTFO_EDGE_INFLOW, IBTFO_EDGE_PNGGEO};
    static const char* pn[] = {"FP64", "mixed"};
    static const int pv[] = {IBTFO_FP64, IBTFO_MIXED};
    int rc;
    if ((rc = lookup(str, "Scheme", sn, sv, 4, &c->scheme))) return rc;
    if ((rc = lookup(str, "Riemann", rn, rv, 4, &c->riemann))) return rc;
    if ((rc = lookup(str, "Limiter", ln, lv, 3, &c->limiter))) return rc;
    if ((rc = lookup(str, "Boundary", bn, bv, 9, &c->boundary))) return rc;
    if ((rc = lookup(str, "Time", tn, tv, 2, &c->time))) return rc;
    if ((rc = lookup(str, "Edge", en, ev, 4, &c->edge))) return rc;
    if ((rc = lookup(str, "Precision", pn, pv, 2, &c->precision))) return rc;
    c->quirks = int(get("quirks", c->quirks));
    if (c->edge == IBTFO_EDGE_PNGGEO) {
        // Test/PNGGeo/Global_Boudary.cpp:61-66: the post-shock state of a Mach 1.3
        // shock into rho = 1, p = 0.71429, in the hook's operation order
        const double gamma = c->gamma, rho1 = 1.0, p1 = 0.71429, Ms = 1.3;
        const double rho2 = rho1 * (gamma + 1) * Ms * Ms / (2 + (gamma - 1) * Ms * Ms);
        const double p2 = p1 * ((2 * gamma) * Ms * Ms / (gamma + 1) - (gamma - 1) / (gamma + 1));
        const double u2 = Ms * (1 - rho1 / rho2);
        c->edge_state[0] = rho2; c->edge_state[1] = u2; c->edge_state[2] = 0.0; c->edge_state[3] = p2;
    }
    c->edge_state[0] = get("inflow_rho", c->edge_state[0]);
    c->edge_state[1] = get("inflow_u", c->edge_state[1]);
    c->edge_state[2] = get("inflow_v", c->edge_state[2]);
    c->edge_state[3] = get("inflow_p", c->edge_state[3]);
    return IBTFO_OK;
}

extern "C" int ibtfo_inp_number(const char* path, const char* key, double* value) {
    std::map<std::string, double> num;
    if (!path || !key || !value) return -IBTFO_E_ARG;
    if (!inp_numbers(path, num)) return -IBTFO_E_IO;
    auto it = num.find(key);
    if (it == num.end()) return 0;
    *value = it->second;
    return 1;
}

extern "C" int ibtfo_inp_string(const char* path, const char* key, char* buf, int cap) {
    std::map<std::string, std::string> str;
    if (!path || !key || !buf || cap <= 0) return -IBTFO_E_ARG;
    if (!inp_strings(path, str)) return -IBTFO_E_IO;
    auto it = str.find(key);
    if (it == str.end()) return 0;
    int n = (int)it->second.size();
    if (n + 1 > cap) return -IBTFO_E_ARG;
    memcpy(buf, it->second.c_str(), n + 1);
    return n;
}
