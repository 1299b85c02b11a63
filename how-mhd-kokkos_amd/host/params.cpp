// Input-file parser of the host shell.
//
// Restates read_parameters / finalize_parameters_after_read
// (src/parameters.cpp:7-160): whitespace-separated key/value pairs read with
// operator>>, a token starting with '#' comments out the rest of its line,
// unknown keys warn on stderr and skip the rest of their line, and nx/ny/nz are
// the global active-cell counts.  Keys this implementation adds: precision,
// indicator, indicator_weights, seed, quiet, checkpoint_every, restart,
// max_steps ("stop after max_steps hydro steps").  nstep is parsed and, as in
// the reference, not used.
#include "howmhd_host.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

namespace {

void set_str(char* dst, size_t cap, const std::string& v) {
    std::snprintf(dst, cap, "%s", v.c_str());
}

void fail(char* errbuf, int errlen, const std::string& msg) {
    if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", msg.c_str());
}

void defaults(howmhd_run* r) {
    std::memset(r, 0, sizeof(*r));
    r->xsize = r->ysize = r->zsize = 1.0;
    r->gam = 5.0 / 3.0;
    r->cour = 0.4;
    r->tend = 0.0;
    r->rhomin = 1.0e-12;
    r->pgmin = 1.0e-12;
    set_str(r->x1bc, sizeof(r->x1bc), "periodic");
    set_str(r->x2bc, sizeof(r->x2bc), "periodic");
    set_str(r->x3bc, sizeof(r->x3bc), "periodic");
    set_str(r->output_format, sizeof(r->output_format), "dat");
    set_str(r->precision, sizeof(r->precision), "fp64");
    set_str(r->indicator, sizeof(r->indicator), "off");
    r->seed = 20240611ULL;
}

}  // namespace

extern "C" int howmhd_host_parse(const char* text, howmhd_run* r, char* errbuf, int errlen) {
    if (!text || !r) {
        fail(errbuf, errlen, "howmhd_host_parse: null argument");
        return HOWMHD_EINVAL;
    }
    defaults(r);
    std::istringstream in(text);
    std::string key, s;
    while (in >> key) {
        if (key.empty()) continue;
        if (key[0] == '#') {
            std::string dummy;
            std::getline(in, dummy);
        } else if (key == "nx" || key == "gnx") in >> r->gnx;
        else if (key == "ny" || key == "gny") in >> r->gny;
        else if (key == "nz" || key == "gnz") in >> r->gnz;
        else if (key == "nt") in >> r->nt;
        else if (key == "nstep") in >> r->nstep;
        else if (key == "ntime") in >> r->ntime;
        else if (key == "xsize") in >> r->xsize;
        else if (key == "ysize") in >> r->ysize;
        else if (key == "zsize") in >> r->zsize;
        else if (key == "gam") in >> r->gam;
        else if (key == "cour") in >> r->cour;
        else if (key == "tend") in >> r->tend;
        else if (key == "rhomin") in >> r->rhomin;
        else if (key == "pgmin") in >> r->pgmin;
        else if (key == "x1bc") { in >> s; set_str(r->x1bc, sizeof(r->x1bc), s); }
        else if (key == "x2bc") { in >> s; set_str(r->x2bc, sizeof(r->x2bc), s); }
        else if (key == "x3bc") { in >> s; set_str(r->x3bc, sizeof(r->x3bc), s); }
        else if (key == "problem") { in >> s; set_str(r->problem, sizeof(r->problem), s); }
        else if (key == "output_format") { in >> s; set_str(r->output_format, sizeof(r->output_format), s); }
        else if (key == "px") in >> r->px;
        else if (key == "py") in >> r->py;
        else if (key == "pz") in >> r->pz;
        else if (key == "precision") { in >> s; set_str(r->precision, sizeof(r->precision), s); }
        else if (key == "indicator") { in >> s; set_str(r->indicator, sizeof(r->indicator), s); }
        else if (key == "indicator_weights") { in >> s; set_str(r->indicator_weights, sizeof(r->indicator_weights), s); }
        else if (key == "seed") { unsigned long long v = 0; in >> v; r->seed = v; }
        else if (key == "quiet") in >> r->quiet;
        else if (key == "checkpoint_every") in >> r->checkpoint_every;
        else if (key == "restart") { in >> s; set_str(r->restart, sizeof(r->restart), s); }
        else if (key == "max_steps") in >> r->max_steps;
        else {
            std::cerr << "Warning: unknown parameter " << key << "\n";
            std::string dummy;
            std::getline(in, dummy);
        }
    }

    // finalize_parameters_after_read (src/parameters.cpp:120-160)
    if (r->gnx <= 0 || r->gny <= 0 || r->gnz <= 0) {
        std::ostringstream m;
        m << "Error: global grid size must be positive.\ngnx = " << r->gnx << " gny = " << r->gny
          << " gnz = " << r->gnz;
        fail(errbuf, errlen, m.str());
        return HOWMHD_EINVAL;
    }
    if (r->nt <= 0) {
        fail(errbuf, errlen, "Error: nt must be positive. nt = " + std::to_string(r->nt));
        return HOWMHD_EINVAL;
    }
    if (r->tend <= 0.0) {
        std::ostringstream m;
        m << "Error: tend must be positive. tend = " << r->tend;
        fail(errbuf, errlen, m.str());
        return HOWMHD_EINVAL;
    }
    r->nx = r->gnx;
    r->ny = r->gny;
    r->nz = r->gnz;
    r->dx = r->xsize / static_cast<double>(r->gnx);
    r->dy = r->ysize / static_cast<double>(r->gny);
    r->dz = r->zsize / static_cast<double>(r->gnz);
    return HOWMHD_OK;
}

namespace {
int bc_code(const char* s, int* out) {
    if (std::strcmp(s, "periodic") == 0) { *out = HOWMHD_BC_PERIODIC; return 0; }
    if (std::strcmp(s, "open") == 0) { *out = HOWMHD_BC_OPEN; return 0; }
    return -1;
}
}  // namespace

extern "C" int howmhd_host_gpu_params(const howmhd_run* r, howmhd_params* p, char* errbuf, int errlen) {
    std::memset(p, 0, sizeof(*p));
    p->nx = r->nx; p->ny = r->ny; p->nz = r->nz;
    p->dx = r->dx; p->dy = r->dy; p->dz = r->dz;
    p->gam = r->gam; p->cour = r->cour; p->rhomin = r->rhomin; p->pgmin = r->pgmin;
    const char* names[3] = {r->x1bc, r->x2bc, r->x3bc};
    const char* keys[3] = {"x1bc", "x2bc", "x3bc"};
    for (int d = 0; d < 3; ++d) {
        if (bc_code(names[d], &p->bc[d]) != 0) {   // src/bound.cpp:305-315
            fail(errbuf, errlen, std::string("Error: unknown ") + keys[d] + " = " + names[d]);
            return HOWMHD_EINVAL;
        }
    }
    if (std::strcmp(r->precision, "fp64") == 0) p->precision = HOWMHD_PREC_FP64;
    else if (std::strcmp(r->precision, "fp32") == 0) p->precision = HOWMHD_PREC_FP32;
    else if (std::strcmp(r->precision, "fp64fast") == 0) p->precision = HOWMHD_PREC_FP64_FAST;
    else {
        fail(errbuf, errlen, std::string("Error: unknown precision = ") + r->precision + " (fp64, fp32, fp64fast)");
        return HOWMHD_EINVAL;
    }
    if (std::strcmp(r->indicator, "off") == 0) p->indicator = HOWMHD_INDICATOR_OFF;
    else if (std::strcmp(r->indicator, "mlp") == 0) p->indicator = HOWMHD_INDICATOR_MLP;
    else {
        fail(errbuf, errlen, std::string("Error: unknown indicator = ") + r->indicator + " (off, mlp)");
        return HOWMHD_EINVAL;
    }
    return HOWMHD_OK;
}

extern "C" int howmhd_host_load_indicator(const char* path, float* params, int nparam, float* threshold,
                                          char* errbuf, int errlen) {
    if (!path || !params || !threshold) {
        fail(errbuf, errlen, "howmhd_host_load_indicator: null argument");
        return HOWMHD_EINVAL;
    }
    std::ifstream f(path);
    if (!f) {
        fail(errbuf, errlen, std::string("Error: cannot open indicator_weights file: ") + path);
        return HOWMHD_EINVAL;
    }
    std::string line;
    bool header = false;
    int count = 0;
    while (std::getline(f, line)) {
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream in(line.substr(b));
        if (!header) {
            std::string tag;
            int n = 0;
            double thr = 0.0;
            if (!(in >> tag >> n >> thr) || tag != "howmhd-mlp") {
                fail(errbuf, errlen, std::string("Error: ") + path + ": header 'howmhd-mlp <nparam> <threshold>' expected");
                return HOWMHD_EINVAL;
            }
            if (n != nparam) {
                fail(errbuf, errlen, std::string("Error: ") + path + ": " + std::to_string(n) + " parameters, expected " +
                                         std::to_string(nparam));
                return HOWMHD_EINVAL;
            }
            if (!(thr > 0.0 && thr < 1.0)) {
                fail(errbuf, errlen, std::string("Error: ") + path + ": threshold must lie in (0, 1)");
                return HOWMHD_EINVAL;
            }
            *threshold = (float)thr;
            header = true;
            continue;
        }
        // one value per line: anything after it is an error too
        double v = 0.0;
        std::string rest;
        if (!(in >> v) || !std::isfinite(v) || (in >> rest) || count >= nparam) {
            fail(errbuf, errlen, std::string("Error: ") + path + ": bad or extra value: " + line);
            return HOWMHD_EINVAL;
        }
        params[count++] = (float)v;
    }
    if (!header || count != nparam) {
        fail(errbuf, errlen, std::string("Error: ") + path + ": " + std::to_string(count) + " of " +
                                 std::to_string(nparam) + " parameters read");
        return HOWMHD_EINVAL;
    }
    return HOWMHD_OK;
}
