// Problem registry and initial conditions of the host shell.
//
// The registry restates src/problem.cpp:7-37 (name -> initialiser; an unknown
// name fails and lists the names).  The six reference problems are restated
// from src/problems/*.cpp with the same expressions in the same order, so the
// host libm and the same rounding give the same q(0).  Three problems are
// added for the benchmark configurations: cpaw3d (oblique circularly polarised
// Alfven wave), mhd_blast3d (blast with p_in=10, p_out=0.1,
// B=(1/sqrt2,1/sqrt2,0)) and turbulence3d (seeded solenoidal velocity field).
#include "howmhd_host.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

struct Grid {
    const howmhd_run* r;
    double* q;
    size_t n6x, n6y, n6z, ncell;
    Grid(const howmhd_run* run, double* q0)
        : r(run), q(q0), n6x((size_t)run->nx + 6), n6y((size_t)run->ny + 6), n6z((size_t)run->nz + 6),
          ncell(n6x * n6y * n6z) {}
    double& at(int m, int i, int j, int k) const {
        return q[(size_t)m * ncell + ((size_t)i * n6y + (size_t)j) * n6z + (size_t)k];
    }
    void set(int i, int j, int k, double rho, double Mx, double My, double Mz, double Bx, double By,
             double Bz, double E) const {
        at(0, i, j, k) = rho; at(1, i, j, k) = Mx; at(2, i, j, k) = My; at(3, i, j, k) = Mz;
        at(4, i, j, k) = Bx; at(5, i, j, k) = By; at(6, i, j, k) = Bz; at(7, i, j, k) = E;
    }
};

using InitFn = void (*)(const Grid&);

// src/problems/uniform.cpp:5-28
void init_uniform(const Grid& g) {
    for (int i = 3; i < g.r->nx + 3; ++i)
        for (int j = 3; j < g.r->ny + 3; ++j)
            for (int k = 3; k < g.r->nz + 3; ++k) g.set(i, j, k, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0);
}

// src/problems/orszag_tang.cpp:6-77
void init_orszag_tang(const Grid& g) {
    const double pi = 3.141592653589793238462643383279502884;
    const howmhd_run& par = *g.r;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const double x = (static_cast<double>(i - 3) + 0.5) * par.dx;
                const double y = (static_cast<double>(j - 3) + 0.5) * par.dy;
                const double rho = 25.0 / (36.0 * pi);
                const double pg = 5.0 / (12.0 * pi);
                const double vx = -std::sin(2.0 * pi * y);
                const double vy = std::sin(2.0 * pi * x);
                const double vz = 0.0;
                const double normB = std::sqrt(4.0 * pi);
                const double Bx = -std::sin(2.0 * pi * y) / normB;
                const double By = std::sin(4.0 * pi * x) / normB;
                const double Bz = 0.0;
                const double Mx = rho * vx, My = rho * vy, Mz = rho * vz;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double E = pg / (par.gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, Mx, My, Mz, Bx, By, Bz, E);
            }
}

// src/problems/brio_wu.cpp:5-66
void init_brio_wu(const Grid& g) {
    const howmhd_run& par = *g.r;
    const double xmid = 0.5 * par.xsize;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const double x = (static_cast<double>(i - 3) + 0.5) * par.dx;
                double rho, vx, vy, vz, pg, Bx, By, Bz;
                if (x < xmid) {
                    rho = 1.0; vx = 0.0; vy = 0.0; vz = 0.0; pg = 1.0; Bx = 0.75; By = 1.0; Bz = 0.0;
                } else {
                    rho = 0.125; vx = 0.0; vy = 0.0; vz = 0.0; pg = 0.1; Bx = 0.75; By = -1.0; Bz = 0.0;
                }
                const double Mx = rho * vx, My = rho * vy, Mz = rho * vz;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double E = pg / (par.gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, Mx, My, Mz, Bx, By, Bz, E);
            }
}

// src/problems/mhd_rotor.cpp:6-111
void init_mhd_rotor(const Grid& g) {
    const double pi = 3.141592653589793238462643383279502884;
    const howmhd_run& par = *g.r;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const double x = (static_cast<double>(i - 3) + 0.5) * par.dx;
                const double y = (static_cast<double>(j - 3) + 0.5) * par.dy;
                const double xc = 0.5, yc = 0.5;
                const double dx = x - xc, dy = y - yc;
                const double r = std::sqrt(dx * dx + dy * dy);
                const double r0 = 0.10, r1 = 0.115;
                const double rho_in = 10.0, rho_out = 1.0;
                const double omega = 20.0;
                double rho, vx, vy;
                if (r <= r0) {
                    rho = rho_in;
                    vx = -omega * dy;
                    vy = omega * dx;
                } else if (r <= r1) {
                    const double f = (r1 - r) / (r1 - r0);
                    rho = rho_out + f * (rho_in - rho_out);
                    vx = f * (-omega * dy);
                    vy = f * (omega * dx);
                } else {
                    rho = rho_out;
                    vx = 0.0;
                    vy = 0.0;
                }
                const double vz = 0.0;
                const double pg = 1.0;
                const double normB = std::sqrt(4.0 * pi);
                const double Bx = 5.0 / normB, By = 0.0, Bz = 0.0;
                const double Mx = rho * vx, My = rho * vy, Mz = rho * vz;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double E = pg / (par.gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, Mx, My, Mz, Bx, By, Bz, E);
            }
}

// Blast body shared by blast3d (src/problems/blast3d.cpp:8-96, constants
// 112-120) and mhd_blast3d: uniform rho, pressure pg_in inside r0 = 0.1 * the
// shortest box side around the box centre, uniform B.
void blast_body(const Grid& g, double pg_out, double pg_in, double Bx0, double By0, double Bz0) {
    const howmhd_run& par = *g.r;
    const double gam = par.gam;
    const double dx = par.dx, dy = par.dy, dz = par.dz;
    const double Lx = par.xsize, Ly = par.ysize, Lz = par.zsize;
    const double xc = 0.5 * Lx, yc = 0.5 * Ly, zc = 0.5 * Lz;
    const double r0 = 0.10 * std::fmin(Lx, std::fmin(Ly, Lz));
    const double rho0 = 1.0;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const int ig = i - 3, jg = j - 3, kg = k - 3;   // one rank: istart = 0
                const double x = (static_cast<double>(ig) + 0.5) * dx;
                const double y = (static_cast<double>(jg) + 0.5) * dy;
                const double z = (static_cast<double>(kg) + 0.5) * dz;
                const double rx = x - xc, ry = y - yc, rz = z - zc;
                const double rr = std::sqrt(rx * rx + ry * ry + rz * rz);
                const double rho = rho0;
                const double vx = 0.0, vy = 0.0, vz = 0.0;
                const double Bx = Bx0, By = By0, Bz = Bz0;
                const double pg = (rr <= r0) ? pg_in : pg_out;
                const double Mx = rho * vx, My = rho * vy, Mz = rho * vz;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double E = pg / (gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, Mx, My, Mz, Bx, By, Bz, E);
            }
}

void init_blast3d(const Grid& g) { blast_body(g, 1.0, 100.0, 0.1, 0.0, 0.0); }

void init_mhd_blast3d(const Grid& g) {
    const double b = 1.0 / std::sqrt(2.0);
    blast_body(g, 0.1, 10.0, b, b, 0.0);
}

// src/problems/mpi_smooth3d.cpp:8-95
void init_mpi_smooth3d(const Grid& g) {
    const double pi = 3.141592653589793238462643383279502884;
    const howmhd_run& par = *g.r;
    const double gam = par.gam;
    const double dx = par.dx, dy = par.dy, dz = par.dz;
    const double Lx = par.xsize, Ly = par.ysize, Lz = par.zsize;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const int ig = i - 3, jg = j - 3, kg = k - 3;
                const double x = (static_cast<double>(ig) + 0.5) * dx;
                const double y = (static_cast<double>(jg) + 0.5) * dy;
                const double z = (static_cast<double>(kg) + 0.5) * dz;
                const double sx = std::sin(2.0 * pi * x / Lx);
                const double sy = std::sin(2.0 * pi * y / Ly);
                const double sz = std::sin(2.0 * pi * z / Lz);
                const double cx = std::cos(2.0 * pi * x / Lx);
                const double cy = std::cos(2.0 * pi * y / Ly);
                const double cz = std::cos(2.0 * pi * z / Lz);
                const double rho = 1.0 + 0.05 * sx * sy * sz;
                const double vx = 0.10 + 0.01 * cy * cz;
                const double vy = 0.05 + 0.01 * cz * cx;
                const double vz = 0.02 + 0.01 * cx * cy;
                const double Bx = 0.10 + 0.005 * sy * sz;
                const double By = 0.08 + 0.005 * sz * sx;
                const double Bz = 0.06 + 0.005 * sx * sy;
                const double pg = 1.0 + 0.02 * cx * cy * cz;
                const double Mx = rho * vx, My = rho * vy, Mz = rho * vz;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double E = pg / (gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, Mx, My, Mz, Bx, By, Bz, E);
            }
}

// Circularly polarised Alfven wave travelling along e1 = (1,2,2)/3 with
// wavelength 1: in a box 3 x 1.5 x 1.5 the phase 2*pi*(x+2y+2z)/3 is periodic.
// rho = 1, p = 0.1, B_par = 1, |B_perp| = 0.1, v_perp = -B_perp, v_par = 0, so
// the Alfven speed is 1 and one period is t = 1.  e2 = (0,1,-1)/sqrt2,
// e3 = e1 x e2 = (-4,1,1)/(3 sqrt2).
void init_cpaw3d(const Grid& g) {
    const double pi = 3.141592653589793238462643383279502884;
    const howmhd_run& par = *g.r;
    const double s2 = std::sqrt(2.0);
    const double e1[3] = {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0};
    const double e2[3] = {0.0, 1.0 / s2, -1.0 / s2};
    const double e3[3] = {-4.0 / (3.0 * s2), 1.0 / (3.0 * s2), 1.0 / (3.0 * s2)};
    const double rho = 1.0, pg = 0.1, bpar = 1.0, amp = 0.1;
    for (int i = 3; i < par.nx + 3; ++i)
        for (int j = 3; j < par.ny + 3; ++j)
            for (int k = 3; k < par.nz + 3; ++k) {
                const double x = (static_cast<double>(i - 3) + 0.5) * par.dx;
                const double y = (static_cast<double>(j - 3) + 0.5) * par.dy;
                const double z = (static_cast<double>(k - 3) + 0.5) * par.dz;
                const double xi = e1[0] * x + e1[1] * y + e1[2] * z;
                const double b2 = amp * std::sin(2.0 * pi * xi);
                const double b3 = amp * std::cos(2.0 * pi * xi);
                double B[3], v[3];
                for (int d = 0; d < 3; ++d) {
                    B[d] = bpar * e1[d] + b2 * e2[d] + b3 * e3[d];
                    v[d] = -(b2 * e2[d] + b3 * e3[d]);
                }
                const double vv2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                const double BB2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
                const double E = pg / (par.gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, rho * v[0], rho * v[1], rho * v[2], B[0], B[1], B[2], E);
            }
}

// SplitMix64 (Steele, Lea & Flood 2014): seeds the turbulence phases.
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Decaying-turbulence start: rho = 1, sonic Mach 1 at v_rms = 1 (p = 1/gam),
// plasma beta 1 with a mean field along z (Bz = sqrt(2p)), and a solenoidal
// velocity field v(x) = sum_k a_k cos(2 pi k.x/L + phi_k) over integer wave
// vectors with 1 <= |k| <= 4 (one of each +-k pair), |a_k| proportional to
// |k|^-1 (so |v_k|^2 ~ k^-2), a_k perpendicular to k with a random direction,
// phases from SplitMix64(seed); finally scaled to v_rms = 1 over the grid.
void init_turbulence3d(const Grid& g) {
    const double pi = 3.141592653589793238462643383279502884;
    const howmhd_run& par = *g.r;
    struct Mode { int kx, ky, kz; double ax, ay, az, ph; };
    std::vector<Mode> modes;
    SplitMix64 rng{par.seed};
    for (int kx = -4; kx <= 4; ++kx)
        for (int ky = -4; ky <= 4; ++ky)
            for (int kz = -4; kz <= 4; ++kz) {
                const int k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < 1 || k2 > 16) continue;
                // keep one of each +-k pair
                if (kx < 0 || (kx == 0 && ky < 0) || (kx == 0 && ky == 0 && kz < 0)) continue;
                const double kn = std::sqrt((double)k2);
                // random direction, made perpendicular to k
                double d[3] = {2.0 * rng.uniform() - 1.0, 2.0 * rng.uniform() - 1.0, 2.0 * rng.uniform() - 1.0};
                const double kd = (d[0] * kx + d[1] * ky + d[2] * kz) / (double)k2;
                d[0] -= kd * kx; d[1] -= kd * ky; d[2] -= kd * kz;
                const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                const double ph = 2.0 * pi * rng.uniform();
                if (dn < 1e-12) continue;
                const double a = 1.0 / kn / dn;
                modes.push_back({kx, ky, kz, a * d[0], a * d[1], a * d[2], ph});
            }
    const int nx = par.nx, ny = par.ny, nz = par.nz;
    std::vector<double> vx((size_t)nx * ny * nz), vy(vx.size()), vz(vx.size());
    // separable evaluation: cos(ax + by + cz + ph) from per-axis phase tables
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (const Mode& md : modes) {
            const double pz = 2.0 * pi * md.kz * ((k + 0.5) * par.dz) / par.zsize + md.ph;
            for (int j = 0; j < ny; ++j) {
                const double pyz = pz + 2.0 * pi * md.ky * ((j + 0.5) * par.dy) / par.ysize;
                for (int i = 0; i < nx; ++i) {
                    const double ph = pyz + 2.0 * pi * md.kx * ((i + 0.5) * par.dx) / par.xsize;
                    const double c = std::cos(ph);
                    const size_t t = ((size_t)k * ny + j) * nx + i;
                    vx[t] += md.ax * c;
                    vy[t] += md.ay * c;
                    vz[t] += md.az * c;
                }
            }
        }
    }
    double s2 = 0.0;
    for (size_t t = 0; t < vx.size(); ++t) s2 += vx[t] * vx[t] + vy[t] * vy[t] + vz[t] * vz[t];
    const double scale = 1.0 / std::sqrt(s2 / (double)vx.size());
    const double rho = 1.0;
    const double pg = 1.0 / par.gam;
    const double Bz0 = std::sqrt(2.0 * pg);
    for (int i = 3; i < nx + 3; ++i)
        for (int j = 3; j < ny + 3; ++j)
            for (int k = 3; k < nz + 3; ++k) {
                const size_t t = ((size_t)(k - 3) * ny + (j - 3)) * nx + (i - 3);
                const double ux = scale * vx[t], uy = scale * vy[t], uz = scale * vz[t];
                const double vv2 = ux * ux + uy * uy + uz * uz;
                const double BB2 = Bz0 * Bz0;
                const double E = pg / (par.gam - 1.0) + 0.5 * rho * vv2 + 0.5 * BB2;
                g.set(i, j, k, rho, rho * ux, rho * uy, rho * uz, 0.0, 0.0, Bz0, E);
            }
}

const std::map<std::string, InitFn>& registry() {
    static const std::map<std::string, InitFn> reg = {
        {"uniform", init_uniform},
        {"orszag_tang", init_orszag_tang},
        {"brio_wu", init_brio_wu},
        {"mhd_rotor", init_mhd_rotor},
        {"blast3d", init_blast3d},
        {"mpi_smooth3d", init_mpi_smooth3d},
        {"cpaw3d", init_cpaw3d},
        {"mhd_blast3d", init_mhd_blast3d},
        {"turbulence3d", init_turbulence3d},
    };
    return reg;
}

}  // namespace

extern "C" const char* howmhd_host_problem_names(void) {
    static std::string names;
    if (names.empty())
        for (const auto& kv : registry()) names += kv.first + "\n";
    return names.c_str();
}

extern "C" int howmhd_host_init_problem(const howmhd_run* run, double* q0, char* errbuf, int errlen) {
    const auto& reg = registry();
    auto it = reg.find(run->problem);
    if (it == reg.end()) {
        std::string msg = std::string("Error: unknown problem = ") + run->problem + "\nAvailable problems:\n";
        for (const auto& kv : reg) msg += "  " + kv.first + "\n";
        if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", msg.c_str());
        return HOWMHD_EINVAL;
    }
    it->second(Grid(run, q0));
    return HOWMHD_OK;
}
