// Device context of one run and the host-side helpers shared by the
// translation units of the GPU library.
//
// The library is built from translation units with different floating-point
// contraction: the plain FP64 form (reference rounding, -ffp-contract=off,
// HOWMHD_EXACT=1) and the FMA forms (fp64fast, fp32; -ffp-contract=fast,
// HOWMHD_EXACT=0).  Each unit instantiates its own kernels (their template
// argument EX equals the unit's HOWMHD_EXACT), and exposes its stage launcher
// through the functions declared at the end of this header.
#pragma once

#include "howmhd_gpu.h"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

struct howmhd_gpu_ctx {
    howmhd_params p{};
    howmhd::Geom g{};
    howmhd::Phys ph{};
    howmhd::GasConst gc{};
    howmhd::CtFlags ctf{};
    size_t ncell = 0;
    size_t rsize = 8;          // bytes of the sweep/CT real type
    // tiles per block of the marching y/z sweeps when positive (environment
    // variable HOWMHD_SWEEP_MARCH at create, for tests of the march on small
    // grids); 0: chosen from the launch size
    int march_force = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double* qbuf[5] = {};      // 8 x ncell each
    double* fbuf[5] = {};      // 3 x ncell each
    int qid[6] = {0, 1, 2, 3, 4, 1};  // stage -> buffer
    void* dF[3] = {};
    void* ct[6] = {};          // fsy fsz gsx gsz hsx hsy
    void* O2[3] = {};
    void* O[3] = {};
    unsigned long long* dmax = nullptr;
    // CFL maxima per block of the last stage-5 combination (3 per block of
    // k_combine); cfl_fresh: they belong to the current q(0)
    double* cflpart = nullptr;
    bool cfl_fresh = false;
    // cells whose density / pressure floor the stage-5 prot applied, per
    // block of the last step's stage-5 combination (2 per block);
    // floors_fresh: they belong to the current q(0)
    unsigned* floorpart = nullptr;
    bool floors_fresh = false;
    // diagnostics: per-block partials of k_diag and the folded values
    double* diagpart = nullptr;
    double* diagout = nullptr;
    // host transfers: one variable in the reference layout
    // ((nx+6)(ny+6)(nz+6) doubles), allocated with the context
    double* stage_ref = nullptr;
    // learned indicator: device weights, routing mode (0 off, 1 on with the
    // plain-WENO fallback, 2 on for every step), the routed sweeps' counters
    // per wavefront slot (nwave slots, cleared by each fold) and their
    // totals (flagged, total, wavefronts flagged, wavefronts total)
    howmhd::MlpDev* mlp = nullptr;
    int ind_mode = 0;
    bool route_now = false;              // the running step routes
    uint4* iwave = nullptr;
    long long nwave = 0;
    // reference-rounding sweeps: tiles to recompute with IEEE divisions
    // (nwave / 4 flags and queue entries, kept zero/empty between sweeps;
    // queue length and blocks done)
    unsigned* fixflag = nullptr;
    unsigned* fixlist = nullptr;
    unsigned* fixctl = nullptr;
    unsigned long long* istats = nullptr;
    // (flagged, total) of the most recent routed step, and whether one ran
    unsigned long long* ilast = nullptr;
    bool any_routed = false;
    // mode 1: counts of the last routed step travel to pinned host memory
    // behind an event; a step whose wavefronts were mostly flagged sends the
    // following steps to the plain sweep
    unsigned long long* istats_host = nullptr;
    hipEvent_t istats_ev = nullptr;
    bool istats_pending = false;
    int fallback_left = 0;
    long long acc[4] = {0, 0, 0, 0};     // flagged, total, waves flagged, waves total
    long long routed_steps = 0, fallback_steps = 0;
    bool faces_ready = false;
    bool timing = false;
    struct Timed {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Timed> timed;
    double tms[HOWMHD_K_COUNT] = {};
    int tcount[HOWMHD_K_COUNT] = {};
    std::string err = "no error";
};

// A failed HIP call is reported once: the runtime's last-error state is
// cleared, so that the next launch check of this thread does not report it
// again.
#define HCHECK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (void)hipGetLastError();                                                       \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return HOWMHD_EHIP;                                                            \
        }                                                                                  \
    } while (0)

namespace howmhd {


inline int fail(howmhd_gpu_ctx* c, int code, const std::string& msg) {
    c->err = msg;
    return code;
}

// Brackets one launch with events when timing is on.
struct TimedLaunch {
    howmhd_gpu_ctx* c;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(howmhd_gpu_ctx* ctx, int k) : c(ctx), kind(k) {
        if (c->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, c->stream);
    }
    ~TimedLaunch() {
        if (a && b) {
            (void)hipEventRecord(b, c->stream);
            c->timed.push_back({kind, a, b});
        }
    }
};

// One SSPRK stage (sweeps, EMFs, faces, combination) of each form.
int stage_fp64(howmhd_gpu_ctx* c, int iw, double dt);      // step_exact.hip
int stage_fp64fast(howmhd_gpu_ctx* c, int iw, double dt);  // step_fast.hip
int stage_fp32(howmhd_gpu_ctx* c, int iw, double dt);      // step_fast.hip
// Indicator samples (X, y) and/or the in-sweep network's logits of the
// stage-0 state along axis dir from the plain FP64 sweep (step_exact.hip);
// null outputs are skipped.
int collect_fp64(howmhd_gpu_ctx* c, int dir, float* X, float* y, float* logits);

}  // namespace howmhd
