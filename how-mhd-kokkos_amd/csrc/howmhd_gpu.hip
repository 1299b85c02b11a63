// C-ABI of the HIP time step (declared in include/howmhd_gpu.h and
// include/howmhd_mlp.h).
//
// One context owns the device state of a run: five stage buffers of q and of
// the faces (stage 5 reuses the buffer of stage 1, which is dead after stage
// 2; the step ends by swapping it with stage 0 instead of the reference's
// copy_stage(q,0,5)), the three directional interface-flux arrays, the six CT
// flux arrays and the six EMF arrays.  All work is queued on the context's
// stream; tstep and downloads synchronise.
//
// This unit is built with -ffp-contract=off -DHOWMHD_EXACT=1: the CFL
// reduction and the face initialisation are the reference's bit for bit.  The
// stage kernels live in step_exact.hip (plain FP64) and step_fast.hip
// (fp64fast, fp32).
#include "howmhd_gpu.h"
#include "howmhd_mlp.h"
#include "context.hpp"
#include "tstep_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#if !HOWMHD_EXACT
#error "howmhd_gpu.hip is built with HOWMHD_EXACT=1"
#endif

using namespace howmhd;

namespace {

// message of the calling thread's last failing call without a context
// (howmhd_gpu_create, howmhd_gpu_tstep_noalloc)
thread_local std::string g_create_error = "no error";

// Plain-WENO fallback of indicator mode 1: a routed step in which at least
// this fraction of wavefronts ran the full WENO5 sends the next
// kFallbackSteps steps to the plain sweep (no network evaluation), after
// which one step routes again.
constexpr double kFallbackWaveFraction = 0.9;
constexpr int kFallbackSteps = 8;

// What howmhd_gpu_build_kind returns: the target and the rounding of the forms.
constexpr const char* kBuildKind = "gfx950: fp64 reference rounding; fp64fast and fp32 with FMA contraction";

// Folds the counter copy of the last routed step into the running totals
// (wait: block until it has arrived; else only if it has) and takes the
// fallback decision from it.
int fold_indicator_stats(howmhd_gpu_ctx* c, bool wait) {
    if (!c->istats_pending) return HOWMHD_OK;
    if (wait) {
        HCHECK(c, hipEventSynchronize(c->istats_ev));
    } else {
        const hipError_t q = hipEventQuery(c->istats_ev);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // not an error: the copy has not arrived yet
            return HOWMHD_OK;
        }
        HCHECK(c, q);
    }
    long long s[4];
    for (int k = 0; k < 4; ++k) s[k] = (long long)c->istats_host[k];
    for (int k = 0; k < 4; ++k) c->acc[k] += s[k];
    c->istats_pending = false;
    if (s[3] > 0 && (double)s[2] >= kFallbackWaveFraction * (double)s[3]) c->fallback_left = kFallbackSteps;
    return HOWMHD_OK;
}

void drain_timing(howmhd_gpu_ctx* c) {
    for (auto& t : c->timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->tms[t.kind] += ms;
            c->tcount[t.kind] += 1;
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->timed.clear();
}

// fp32 -> bf16 -> fp32, round to nearest even (the device's f2bf)
float bf16_round(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}
uint16_t bf16_bits(float f) {
    const float r = bf16_round(f);
    uint32_t u;
    std::memcpy(&u, &r, 4);
    return (uint16_t)(u >> 16);
}

// Transfers between the caller's arrays in the reference layout
// (m, i, j, k), three ghost layers per side, and the device layout
// (m, z, y, x) of active cells.  One variable at a time goes through the
// context's staging buffer of one variable's reference-layout size: the copy
// reads or writes the caller's (pageable) memory directly and the transpose
// runs on the device (k_ref_to_dev, k_dev_to_ref).
size_t ref_size(const howmhd_gpu_ctx* c) {
    return ((size_t)c->g.n[0] + 6) * ((size_t)c->g.n[1] + 6) * ((size_t)c->g.n[2] + 6);
}

unsigned lay_planes(int n) { return (unsigned)std::min(n, 65535); }

int upload_ref(howmhd_gpu_ctx* c, const double* ref, double* dev, int nvar) {
    const size_t n6 = ref_size(c);
    const dim3 grid(blocks_for(c->g.n[2], LAY_T), blocks_for(c->g.n[0], LAY_T), lay_planes(c->g.n[1]));
    for (int m = 0; m < nvar; ++m) {
        HCHECK(c, hipMemcpyAsync(c->stage_ref, ref + (size_t)m * n6, n6 * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(k_ref_to_dev, grid, dim3(LAY_T, 8), 0, c->stream, (const double*)c->stage_ref,
                           dev + (size_t)m * c->ncell, c->g);
        HCHECK(c, hipGetLastError());
    }
    HCHECK(c, hipStreamSynchronize(c->stream));
    return HOWMHD_OK;
}

int download_ref(howmhd_gpu_ctx* c, const double* dev, double* ref, int nvar) {
    const size_t n6 = ref_size(c);
    const dim3 grid(blocks_for(c->g.n[2] + 6, LAY_T), blocks_for(c->g.n[0] + 6, LAY_T), lay_planes(c->g.n[1] + 6));
    for (int m = 0; m < nvar; ++m) {
        hipLaunchKernelGGL(k_dev_to_ref, grid, dim3(LAY_T, 8), 0, c->stream, dev + (size_t)m * c->ncell,
                           c->stage_ref, c->g);
        HCHECK(c, hipGetLastError());
        HCHECK(c, hipMemcpyAsync(ref + (size_t)m * n6, c->stage_ref, n6 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HCHECK(c, hipStreamSynchronize(c->stream));
    return HOWMHD_OK;
}

// dt of src/tstep.cpp:123-135 from the per-direction maxima
double dt_from_vmax(const howmhd_params& p, const double* vmax) {
    const double tiny = 1.0e-300;
    double invdt = 0.0;
    if (p.nx > 1) invdt += std::fmax(vmax[0], tiny) / p.dx;
    if (p.ny > 1) invdt += std::fmax(vmax[1], tiny) / p.dy;
    if (p.nz > 1) invdt += std::fmax(vmax[2], tiny) / p.dz;
    if (invdt <= 0.0) invdt = std::fmax(vmax[0], tiny) / p.dx;
    return p.cour / invdt;
}

// Largest axis length the launch grids accept on the y and z axes: the
// sweeps, the face update and the EMF build put an axis length (or a tile
// count of it) into gridDim.y or gridDim.z.
constexpr int kMaxYZ = 65535;

// Parameter checks of howmhd_gpu_create (no HIP call); "" when valid.
std::string check_params(const howmhd_params* p, const std::string& who = "howmhd_gpu_create") {
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int d = 0; d < 3; ++d) {
        if (n[d] < 1 || n[d] == 2) return who + ": each axis needs n == 1 or n >= 3 active cells";
        if (p->bc[d] != HOWMHD_BC_PERIODIC && p->bc[d] != HOWMHD_BC_OPEN)
            return who + ": unknown boundary condition code";
    }
    if (p->ny > kMaxYZ || p->nz > kMaxYZ)
        return who + ": ny and nz must not exceed " + std::to_string(kMaxYZ) +
               " (launch grid limit on the y and z axes); got ny = " + std::to_string(p->ny) +
               ", nz = " + std::to_string(p->nz);
    if ((long long)p->nx * p->ny * p->nz > (1LL << 40)) return who + ": more than 2^40 active cells";
    if (p->precision != HOWMHD_PREC_FP64 && p->precision != HOWMHD_PREC_FP32 && p->precision != HOWMHD_PREC_FP64_FAST)
        return who + ": unknown precision";
    if (p->indicator != HOWMHD_INDICATOR_OFF && p->indicator != HOWMHD_INDICATOR_MLP)
        return who + ": unknown indicator";
    if (!(p->dx > 0 && p->dy > 0 && p->dz > 0) || !(p->gam > 1.0))
        return who + ": dx, dy, dz must be > 0 and gam > 1";
    return "";
}

// Scratch of howmhd_gpu_tstep_noalloc: CFL_REF_BLOCKS per-block maximum
// triples and their fold (24.6 KB).  Each GPU the library runs on holds its
// own copy from the first call on it until the process ends.  In host code
// the array's name is not a device address: kernels and copies use the
// address hipGetSymbolAddress returns on the current device.
static __device__ __align__(256) double scratch[3 * CFL_REF_BLOCKS + 3];

constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on device dev (current), looked up once per
// device; nullptr (and *err set) when the runtime refuses.
double* scratch_on(int dev, std::string* err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            *err = std::string("howmhd_gpu_tstep_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_dev[dev] = (double*)ptr;
    }
    return g_scratch_dev[dev];
}

}  // namespace

extern "C" {

int howmhd_gpu_create(const howmhd_params* p, howmhd_gpu_ctx** out) {
    if (!p || !out) {
        g_create_error = "howmhd_gpu_create: null argument";
        return HOWMHD_EINVAL;
    }
    *out = nullptr;
    const std::string bad = check_params(p);
    if (!bad.empty()) {
        g_create_error = bad;
        return HOWMHD_EINVAL;
    }
    const int n[3] = {p->nx, p->ny, p->nz};
    howmhd_gpu_ctx* c = new howmhd_gpu_ctx();
    c->p = *p;
    for (int d = 0; d < 3; ++d) {
        c->g.n[d] = n[d];
        c->g.per[d] = (p->bc[d] == HOWMHD_BC_PERIODIC) ? 1 : 0;
    }
    c->ncell = (size_t)n[0] * n[1] * n[2];
    c->g.ncell = (long long)c->ncell;
    c->ph.gam = p->gam; c->ph.rhomin = p->rhomin; c->ph.pgmin = p->pgmin;
    // IEEE quotients on the host: the bits the device's divisions give
    c->gc.gam2 = (p->gam - 2.0) / (p->gam - 1.0);
    c->gc.gm1 = p->gam - 1.0;
    c->gc.inv_gm1 = 1.0 / c->gc.gm1;
    c->ctf.use_x = n[0] > 1; c->ctf.use_y = n[1] > 1; c->ctf.use_z = n[2] > 1;
    c->ctf.use_Ox = c->ctf.use_x && c->ctf.use_y;
    c->ctf.use_Oy = c->ctf.use_y && c->ctf.use_z;
    c->ctf.use_Oz = c->ctf.use_x && c->ctf.use_z;
    c->ctf.active = c->ctf.use_Ox || c->ctf.use_Oy || c->ctf.use_Oz;
    c->rsize = (p->precision == HOWMHD_PREC_FP32) ? sizeof(float) : sizeof(double);
    if (const char* e = std::getenv("HOWMHD_SWEEP_MARCH")) c->march_force = std::max(0, std::atoi(e));

    auto cleanup_fail = [&](int code, const std::string& msg) {
        (void)hipGetLastError();  // reported here, not at a later launch check
        g_create_error = msg;
        howmhd_gpu_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return cleanup_fail(HOWMHD_EHIP, "hipGetDevice failed (no GPU?)");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(HOWMHD_EHIP, "hipStreamCreate failed");
    // a failed allocation is handled here: the runtime's last error is
    // cleared so that it does not surface at a later launch check.  The
    // clears run on the context's stream: the stream does not wait for the
    // null stream (hipStreamNonBlocking), and a clear there could still be
    // running when the first upload writes the buffer.
    auto alloc = [&](void** ptr, size_t bytes) -> bool {
        if (hipMalloc(ptr, bytes) != hipSuccess || hipMemsetAsync(*ptr, 0, bytes, c->stream) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return true;
    };
    bool ok = true;
    for (int b = 0; b < 5 && ok; ++b) {
        ok = ok && alloc((void**)&c->qbuf[b], 8 * c->ncell * sizeof(double));
        ok = ok && alloc((void**)&c->fbuf[b], 3 * c->ncell * sizeof(double));
    }
    for (int d = 0; d < 3 && ok; ++d) {
        if (n[d] > 1) {
            const size_t nif = c->ncell / n[d] * (size_t)(n[d] + 1);
            ok = ok && alloc(&c->dF[d], 7 * nif * c->rsize);
        }
    }
    for (int t = 0; t < 6 && ok; ++t) ok = ok && alloc(&c->ct[t], c->ncell * c->rsize);
    // the unsmoothed EMFs O2 live in memory only on 1-D/2-D grids; on 3-D
    // grids k_emf_march keeps them on chip
    const bool emf_fused = c->ctf.use_x && c->ctf.use_y && c->ctf.use_z;
    for (int t = 0; t < 3 && ok; ++t) {
        if (!emf_fused) ok = ok && alloc(&c->O2[t], c->ncell * c->rsize);
        ok = ok && alloc(&c->O[t], c->ncell * c->rsize);
    }
    ok = ok && alloc((void**)&c->dmax, 4 * sizeof(unsigned long long));
    ok = ok && alloc((void**)&c->cflpart, 3 * (size_t)blocks_for(c->ncell, 256) * sizeof(double));
    ok = ok && alloc((void**)&c->floorpart, 2 * (size_t)blocks_for(c->ncell, 256) * sizeof(unsigned));
    ok = ok && alloc((void**)&c->diagpart, (size_t)DIAG_BLOCKS * DIAG_NV * sizeof(double));
    ok = ok && alloc((void**)&c->diagout, DIAG_NV * sizeof(double));
    ok = ok && alloc((void**)&c->stage_ref, ref_size(c) * sizeof(double));
    ok = ok && alloc((void**)&c->mlp, sizeof(MlpDev));
    // one counter slot per wavefront of the largest sweep grid
    for (int d = 0; d < 3; ++d) {
        const dim3 gr = sweep_grid(c->g, d);
        c->nwave = std::max(c->nwave, 4LL * gr.x * gr.y * gr.z);
    }
    ok = ok && alloc((void**)&c->iwave, (size_t)c->nwave * sizeof(uint4));
    ok = ok && alloc((void**)&c->fixflag, (size_t)(c->nwave / 4) * sizeof(unsigned));
    ok = ok && alloc((void**)&c->fixlist, (size_t)(c->nwave / 4) * sizeof(unsigned));
    ok = ok && alloc((void**)&c->fixctl, 2 * sizeof(unsigned));
    ok = ok && alloc((void**)&c->istats, 4 * sizeof(unsigned long long));
    ok = ok && alloc((void**)&c->ilast, 2 * sizeof(unsigned long long));
    ok = ok && hipHostMalloc((void**)&c->istats_host, 4 * sizeof(unsigned long long), hipHostMallocDefault) ==
                   hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->istats_ev, hipEventDisableTiming) == hipSuccess;
    if (!ok) return cleanup_fail(HOWMHD_ENOMEM, "howmhd_gpu_create: device allocation failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return cleanup_fail(HOWMHD_EHIP, "howmhd_gpu_create: clearing the device buffers failed");
    *out = c;
    return HOWMHD_OK;
}

int howmhd_gpu_destroy(howmhd_gpu_ctx* c) {
    if (!c) return HOWMHD_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int b = 0; b < 5; ++b) {
        if (c->qbuf[b]) (void)hipFree(c->qbuf[b]);
        if (c->fbuf[b]) (void)hipFree(c->fbuf[b]);
    }
    for (int d = 0; d < 3; ++d) {
        if (c->dF[d]) (void)hipFree(c->dF[d]);
        if (c->O2[d]) (void)hipFree(c->O2[d]);
        if (c->O[d]) (void)hipFree(c->O[d]);
    }
    for (int t = 0; t < 6; ++t)
        if (c->ct[t]) (void)hipFree(c->ct[t]);
    if (c->dmax) (void)hipFree(c->dmax);
    if (c->cflpart) (void)hipFree(c->cflpart);
    if (c->floorpart) (void)hipFree(c->floorpart);
    if (c->diagpart) (void)hipFree(c->diagpart);
    if (c->diagout) (void)hipFree(c->diagout);
    if (c->stage_ref) (void)hipFree(c->stage_ref);
    if (c->ilast) (void)hipFree(c->ilast);
    if (c->mlp) (void)hipFree(c->mlp);
    if (c->istats) (void)hipFree(c->istats);
    if (c->iwave) (void)hipFree(c->iwave);
    if (c->fixflag) (void)hipFree(c->fixflag);
    if (c->fixlist) (void)hipFree(c->fixlist);
    if (c->fixctl) (void)hipFree(c->fixctl);
    if (c->istats_host) (void)hipHostFree(c->istats_host);
    if (c->istats_ev) (void)hipEventDestroy(c->istats_ev);
    drain_timing(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HOWMHD_OK;
}

int howmhd_gpu_upload_q0(howmhd_gpu_ctx* c, const double* q0) {
    if (!c || !q0) return HOWMHD_EINVAL;
    const int rc = upload_ref(c, q0, c->qbuf[c->qid[0]], 8);
    if (rc) return rc;
    c->faces_ready = false;
    c->cfl_fresh = false;
    c->floors_fresh = false;
    return HOWMHD_OK;
}

int howmhd_gpu_init_faces(howmhd_gpu_ctx* c) {
    if (!c) return HOWMHD_EINVAL;
    hipLaunchKernelGGL(k_init_faces, dim3(blocks_for(c->ncell, 256)), dim3(256), 0, c->stream,
                       (const double*)c->qbuf[c->qid[0]], c->fbuf[c->qid[0]], c->g);
    HCHECK(c, hipGetLastError());
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->faces_ready = true;
    return HOWMHD_OK;
}

int howmhd_gpu_upload_faces(howmhd_gpu_ctx* c, const double* b) {
    if (!c || !b) return HOWMHD_EINVAL;
    const int rc = upload_ref(c, b, c->fbuf[c->qid[0]], 3);
    if (rc) return rc;
    c->faces_ready = true;
    return HOWMHD_OK;
}

int howmhd_gpu_tstep(howmhd_gpu_ctx* c, double* dt) {
    if (!c || !dt) return HOWMHD_EINVAL;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->dmax, 3);
    HCHECK(c, hipGetLastError());
    {
        TimedLaunch tl(c, HOWMHD_K_CFL);
        if (c->cfl_fresh) {
            // maxima of the last step's stage-5 combination, which wrote q(0)
            const long long nblk = blocks_for(c->ncell, 256);
            const unsigned nb = std::min<unsigned>(blocks_for((size_t)nblk, 256), 256u);
            hipLaunchKernelGGL(k_cfl_fold, dim3(nb), dim3(256), 0, c->stream, (const double*)c->cflpart, nblk, c->dmax);
        } else {
            const unsigned nb = std::min<unsigned>(blocks_for(c->ncell, 256), 2048u);
            hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, c->stream, (const double*)c->qbuf[c->qid[0]], c->g,
                               c->ph, c->dmax);
        }
    }
    HCHECK(c, hipGetLastError());
    unsigned long long h[3];
    HCHECK(c, hipMemcpyAsync(h, c->dmax, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    double vmax[3];
    for (int d = 0; d < 3; ++d) std::memcpy(&vmax[d], &h[d], sizeof(double));
    *dt = dt_from_vmax(c->p, vmax);
    return HOWMHD_OK;
}

int howmhd_gpu_tstep_noalloc(const howmhd_params* p, const double* q0_dev, void* stream, double* dt) {
    static const char* who = "howmhd_gpu_tstep_noalloc";
    if (!p || !q0_dev || !dt) {
        g_create_error = std::string(who) + ": null argument";
        return HOWMHD_EINVAL;
    }
    const std::string bad = check_params(p, who);
    if (!bad.empty()) {
        g_create_error = bad;
        return HOWMHD_EINVAL;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        g_create_error = std::string(who) + ": no current HIP device (or device index >= 64)";
        return HOWMHD_EHIP;
    }
    std::string err;
    double* part = scratch_on(dev, &err);
    if (!part) {
        g_create_error = err;
        return HOWMHD_EHIP;
    }
    Geom g{};
    g.n[0] = p->nx;
    g.n[1] = p->ny;
    g.n[2] = p->nz;
    g.ncell = (long long)p->nx * p->ny * p->nz;
    Phys ph{p->gam, p->rhomin, p->pgmin};
    const hipStream_t st = (hipStream_t)stream;
    const unsigned nb = (unsigned)std::min<long long>((g.ncell + 255) / 256, CFL_REF_BLOCKS);
    hipLaunchKernelGGL(k_cfl_ref, dim3(nb), dim3(256), 0, st, q0_dev, g, ph, part);
    hipLaunchKernelGGL(k_cfl_ref_fold, dim3(1), dim3(256), 0, st, part, (int)nb);
    double vmax[3];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(vmax, part + 3 * CFL_REF_BLOCKS, sizeof(vmax), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_create_error = std::string(who) + ": " + hipGetErrorString(e);
        return HOWMHD_EHIP;
    }
    *dt = dt_from_vmax(*p, vmax);
    return HOWMHD_OK;
}

int howmhd_gpu_ssprk(howmhd_gpu_ctx* c, double dt) {
    if (!c) return HOWMHD_EINVAL;
    // the reference's loop stops before stepping with dt <= 0
    // (src/main.cpp:175-182); a negative or non-finite dt is refused here
    if (!(dt >= 0.0) || !std::isfinite(dt))
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_ssprk: dt must be finite and >= 0");
    if (!c->faces_ready) return fail(c, HOWMHD_ESTATE, "howmhd_gpu_ssprk: call init_faces or upload_faces first");
    if (c->p.indicator == HOWMHD_INDICATOR_MLP && !c->ind_mode)
        return fail(c, HOWMHD_ESTATE, "howmhd_gpu_ssprk: indicator mlp needs weights (howmhd_gpu_set_indicator)");
    auto stage = (c->p.precision == HOWMHD_PREC_FP32)        ? &stage_fp32
                 : (c->p.precision == HOWMHD_PREC_FP64_FAST) ? &stage_fp64fast
                                                             : &stage_fp64;
    c->route_now = (c->ind_mode == 2);
    if (c->ind_mode == 2) ++c->routed_steps;
    if (c->ind_mode == 1) {
        // the last routed step's counts decide this step (waits for that step)
        const int rc = fold_indicator_stats(c, true);
        if (rc) return rc;
        if (c->fallback_left > 0) {
            --c->fallback_left;
            ++c->fallback_steps;
        } else {
            c->route_now = true;
            ++c->routed_steps;
        }
    }
    c->cfl_fresh = false;
    for (int iw = 1; iw <= 5; ++iw) {
        const int rc = stage(c, iw, dt);
        if (rc) return rc;
    }
    if (c->route_now) {
        // the step's per-wavefront counters into the totals and into the
        // step's own (flagged, total) (slots cleared)
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->ilast, 2);
        HCHECK(c, hipGetLastError());
        const unsigned nb = std::min(1024u, blocks_for((size_t)c->nwave, 256));
        hipLaunchKernelGGL(k_fold_wave_stats, dim3(nb), dim3(256), 0, c->stream, c->iwave, c->nwave, c->istats,
                           c->ilast);
        HCHECK(c, hipGetLastError());
        c->any_routed = true;
    }
    if (c->ind_mode == 1 && c->route_now && !c->istats_pending) {
        // this step's totals to the host, then clear them for the next
        HCHECK(c, hipMemcpyAsync(c->istats_host, c->istats, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 c->stream));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->istats, 4);
        HCHECK(c, hipGetLastError());
        HCHECK(c, hipEventRecord(c->istats_ev, c->stream));
        c->istats_pending = true;
    }
    c->route_now = false;
    // q(0) <- q(5), faces likewise: exchange the buffers of stages 0 and 5
    std::swap(c->qid[0], c->qid[1]);
    c->qid[5] = c->qid[1];
    c->cfl_fresh = true;     // the stage-5 combination left the CFL maxima of the new q(0)
    c->floors_fresh = true;  // and the counts of the floors its prot applied
    return HOWMHD_OK;
}

int howmhd_gpu_download_q0(howmhd_gpu_ctx* c, double* q0) {
    if (!c || !q0) return HOWMHD_EINVAL;
    return download_ref(c, c->qbuf[c->qid[0]], q0, 8);
}

int howmhd_gpu_download_faces(howmhd_gpu_ctx* c, double* b) {
    if (!c || !b) return HOWMHD_EINVAL;
    return download_ref(c, c->fbuf[c->qid[0]], b, 3);
}

int howmhd_gpu_diag(howmhd_gpu_ctx* c, howmhd_diag* d) {
    if (!c || !d) return HOWMHD_EINVAL;
    std::memset(d, 0, sizeof(*d));
    DiagArgs a;
    a.q = c->qbuf[c->qid[0]];
    a.b = c->fbuf[c->qid[0]];
    a.g = c->g;
    a.gam = c->p.gam;
    a.dx = c->p.dx;
    a.dy = c->p.dy;
    a.dz = c->p.dz;
    a.floors = c->floors_fresh ? c->floorpart : nullptr;
    a.nfloor = blocks_for(c->ncell, 256);
    a.part = c->diagpart;
    const int nb = (int)std::min<size_t>(DIAG_BLOCKS, blocks_for(c->ncell, 256));
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, c->stream, a);
    HCHECK(c, hipGetLastError());
    hipLaunchKernelGGL(k_diag_fold, dim3(1), dim3(256), 0, c->stream, (const double*)c->diagpart, nb, c->diagout);
    HCHECK(c, hipGetLastError());
    double h[DIAG_NV];
    unsigned long long il[2] = {0ull, 0ull};
    HCHECK(c, hipMemcpyAsync(h, c->diagout, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    if (c->any_routed) HCHECK(c, hipMemcpyAsync(il, c->ilast, sizeof(il), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < 8; ++m) d->sum_cons[m] = h[DG_SUM + m];
    d->max_divb_ct = h[DG_DIVCT];
    d->max_divb_cell = h[DG_DIVC];
    d->min_rho = -h[DG_NRHO];
    d->min_p = -h[DG_NP];
    d->n_rho_floor = (int64_t)h[DG_NFRHO];
    d->n_p_floor = (int64_t)h[DG_NFP];
    d->flagged_fraction = il[1] ? (double)il[0] / (double)il[1] : 0.0;
    return HOWMHD_OK;
}

int howmhd_gpu_set_timing(howmhd_gpu_ctx* c, int on) {
    if (!c) return HOWMHD_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    drain_timing(c);
    for (int k = 0; k < HOWMHD_K_COUNT; ++k) { c->tms[k] = 0.0; c->tcount[k] = 0; }
    c->timing = (on != 0);
    return HOWMHD_OK;
}

int howmhd_gpu_get_timing(howmhd_gpu_ctx* c, howmhd_timing* t) {
    if (!c || !t) return HOWMHD_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    drain_timing(c);
    for (int k = 0; k < HOWMHD_K_COUNT; ++k) { t->ms[k] = c->tms[k]; t->launches[k] = c->tcount[k]; }
    return HOWMHD_OK;
}

int howmhd_gpu_synchronize(howmhd_gpu_ctx* c) {
    if (!c) return HOWMHD_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    return HOWMHD_OK;
}

const char* howmhd_gpu_last_error(const howmhd_gpu_ctx* c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

int howmhd_gpu_collect_samples(howmhd_gpu_ctx* c, int dir, float* X, float* y, int64_t capacity, int64_t* count) {
    if (!c || !X || !y || !count || dir < 0 || dir > 2) return HOWMHD_EINVAL;
    *count = 0;
    const int n = c->g.n[dir];
    if (n <= 1) return HOWMHD_OK;
    const int64_t need = (int64_t)(c->ncell / n) * (c->g.per[dir] ? n : n + 1);
    if (capacity < need)
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_collect_samples: capacity " + std::to_string(capacity) +
                                          " < " + std::to_string(need) + " interfaces");
    if (c->p.precision != HOWMHD_PREC_FP64)
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_collect_samples: labels come from the plain FP64 solver (precision fp64)");
    // the sweep writes this direction's dF and CT terms too; the next stage
    // recomputes them from its own source
    const int rc = collect_fp64(c, dir, X, y, nullptr);
    if (rc) return rc;
    HCHECK(c, hipStreamSynchronize(c->stream));
    *count = need;
    return HOWMHD_OK;
}

int howmhd_gpu_indicator_logits(howmhd_gpu_ctx* c, int dir, float* logits, int64_t capacity, int64_t* count) {
    if (!c || !logits || !count || dir < 0 || dir > 2) return HOWMHD_EINVAL;
    *count = 0;
    if (!c->ind_mode) return fail(c, HOWMHD_ESTATE, "howmhd_gpu_indicator_logits: no weights loaded");
    const int n = c->g.n[dir];
    if (n <= 1) return HOWMHD_OK;
    const int64_t need = (int64_t)(c->ncell / n) * (c->g.per[dir] ? n : n + 1);
    if (capacity < need)
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_indicator_logits: capacity " + std::to_string(capacity) + " < " +
                                          std::to_string(need) + " interfaces");
    if (c->p.precision != HOWMHD_PREC_FP64)
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_indicator_logits: needs precision fp64");
    const int rc = collect_fp64(c, dir, nullptr, nullptr, logits);
    if (rc) return rc;
    HCHECK(c, hipStreamSynchronize(c->stream));
    *count = need;
    return HOWMHD_OK;
}

int howmhd_gpu_set_indicator(howmhd_gpu_ctx* c, const float* params, float threshold, int enable) {
    if (!c) return HOWMHD_EINVAL;
    if (enable < 0 || enable > 2) return fail(c, HOWMHD_EINVAL, "howmhd_gpu_set_indicator: enable is 0, 1 or 2");
    if (enable == 0) {
        HCHECK(c, hipStreamSynchronize(c->stream));
        c->ind_mode = 0;
        return HOWMHD_OK;
    }
    if (enable && !params) return fail(c, HOWMHD_EINVAL, "howmhd_gpu_set_indicator: null params");
    if (enable && !(threshold > 0.f && threshold < 1.f))
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_set_indicator: threshold must lie in (0, 1)");
    HCHECK(c, hipStreamSynchronize(c->stream));
    if (enable) {
        for (int j = 0; j < HOWMHD_MLP_NPARAM; ++j)
            if (!std::isfinite(params[j]))
                return fail(c, HOWMHD_EINVAL, "howmhd_gpu_set_indicator: non-finite weight " + std::to_string(j));
        MlpDev h;
        std::memset(&h, 0, sizeof(h));
        // A fragments of the sweep's 32x32x16 products (mlp_forward_wave):
        // lane l holds row m = l % 32 and k-slots v of half-wave hw = l / 32
        for (int l = 0; l < 64; ++l) {
            const int m = l % 32, hw = l / 32;
            for (int s = 0; s < 2; ++s)
                for (int v = 0; v < 8; ++v) {
                    // layer 1: k-step 0 slot (hw, v) = feature 8 hw + v; k-step 1 slot (0, v) = 16 + v
                    const int f = s == 0 ? 8 * hw + v : (hw == 0 ? 16 + v : MLP_IN);
                    h.W1f[l][s][v] = f < MLP_IN ? bf16_bits(params[MLP_OFF_W1 + m * MLP_IN + f]) : 0;
                    const int j = mlp_hidden_slot(s, v, hw);
                    h.W2f[l][s][v] = bf16_bits(params[MLP_OFF_W2 + m * MLP_H + j]);
                    h.W3f[l][s][v] = m == 0 ? bf16_bits(params[MLP_OFF_W3 + j]) : 0;
                }
        }
        for (int hw = 0; hw < 2; ++hw)
            for (int r = 0; r < 16; ++r) {
                h.b1r[hw][r] = params[MLP_OFF_B1 + mlp_acc_row(r, hw)];
                h.b2r[hw][r] = params[MLP_OFF_B2 + mlp_acc_row(r, hw)];
            }
        h.b3 = params[MLP_OFF_B3];
        h.logit_threshold = std::log(threshold / (1.f - threshold));
        // on the context's stream (which does not wait for the null stream),
        // complete before h goes out of scope
        HCHECK(c, hipMemcpyAsync(c->mlp, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
        HCHECK(c, hipStreamSynchronize(c->stream));
    }
    HCHECK(c, hipMemsetAsync(c->istats, 0, 4 * sizeof(unsigned long long), c->stream));
    HCHECK(c, hipMemsetAsync(c->ilast, 0, 2 * sizeof(unsigned long long), c->stream));
    c->any_routed = false;
    HCHECK(c, hipMemsetAsync(c->iwave, 0, (size_t)c->nwave * sizeof(uint4), c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->istats_pending = false;
    c->fallback_left = 0;
    for (int k = 0; k < 4; ++k) c->acc[k] = 0;
    c->routed_steps = c->fallback_steps = 0;
    c->ind_mode = enable;
    return HOWMHD_OK;
}

int howmhd_gpu_indicator_stats(howmhd_gpu_ctx* c, int64_t* flagged, int64_t* total, int64_t* waves_flagged,
                               int64_t* waves_total) {
    if (!c || !flagged || !total || !waves_flagged || !waves_total) return HOWMHD_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    const int rc = fold_indicator_stats(c, true);
    if (rc) return rc;
    unsigned long long h[4];
    HCHECK(c, hipMemcpyAsync(h, c->istats, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipMemsetAsync(c->istats, 0, sizeof(h), c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    long long s[4] = {c->acc[0], c->acc[1], c->acc[2], c->acc[3]};
    for (int k = 0; k < 4; ++k) s[k] += (long long)h[k];
    for (int k = 0; k < 4; ++k) c->acc[k] = 0;
    *flagged = s[0];
    *total = s[1];
    *waves_flagged = s[2];
    *waves_total = s[3];
    return HOWMHD_OK;
}

int howmhd_gpu_indicator_routing(howmhd_gpu_ctx* c, int64_t* routed_steps, int64_t* fallback_steps) {
    if (!c || !routed_steps || !fallback_steps) return HOWMHD_EINVAL;
    *routed_steps = c->routed_steps;
    *fallback_steps = c->fallback_steps;
    return HOWMHD_OK;
}

const char* howmhd_gpu_build_kind(void) { return kBuildKind; }

}  // extern "C"
