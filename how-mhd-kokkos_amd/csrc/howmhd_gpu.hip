// C-ABI of the HIP time step (declared in include/howmhd_gpu.h).
//
// One context owns the device state of a run: five stage buffers of q and of
// the faces (stage 5 reuses the buffer of stage 1, which is dead after stage
// 2; the step ends by swapping it with stage 0 instead of the reference's
// copy_stage(q,0,5)), the three directional interface-flux arrays, the six CT
// flux arrays and the six EMF arrays.  All work is queued on the context's
// stream; tstep and downloads synchronise.
#include "howmhd_gpu.h"
#include "howmhd_mlp.h"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace howmhd;

namespace {

std::string g_create_error = "no error";

struct RK {
    double k1, k2, k3, k4, k5;
};

// SSPRK(5,4) coefficients, src/ssprk.cpp:66-94
RK rk_coeff(int iw) {
    RK c{0.0, 0.0, 0.0, 0.0, 0.0};
    switch (iw) {
    case 1: c.k1 = 1.0; c.k2 = 0.39175222700392; c.k3 = 0.0; break;
    case 2: c.k1 = 0.44437049406734; c.k2 = 0.36841059262959; c.k3 = 0.55562950593266; break;
    case 3: c.k1 = 0.62010185138540; c.k2 = 0.25189177424738; c.k3 = 0.37989814861460; break;
    case 4: c.k1 = 0.17807995410773; c.k2 = 0.54497475021237; c.k3 = 0.82192004589227; break;
    case 5:
        c.k1 = -2.081261929715610e-02; c.k2 = 0.22600748319395; c.k3 = 5.03580947213895e-1;
        c.k4 = 0.51723167208978; c.k5 = -6.518979800418380e-12;
        break;
    }
    return c;
}

}  // namespace

struct howmhd_gpu_ctx {
    howmhd_params p{};
    Geom g{};
    Phys ph{};
    CtFlags ctf{};
    size_t ncell = 0;
    size_t rsize = 8;          // bytes of the sweep/CT real type
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
    // learned indicator: device weights, routing mode (0 off, 1 on, 2 on and
    // counted), counter slots
    MlpDev* mlp = nullptr;
    int ind_mode = 0;
    unsigned long long* istats = nullptr;
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

#define HCHECK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return HOWMHD_EHIP;                                                            \
        }                                                                                  \
    } while (0)

namespace {

unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

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

int fail(howmhd_gpu_ctx* c, int code, const std::string& msg) {
    c->err = msg;
    return code;
}

template <typename R, int MODE>
void launch_sweep(howmhd_gpu_ctx* c, int d, const SweepArgs<R>& a) {
    const Geom& g = c->g;
    const int nif = g.per[d] ? g.n[d] : g.n[d] + 1;
    if (d == 0) {
        dim3 grid(blocks_for(nif, SweepTile<0>::LANES), blocks_for(g.n[1], SweepTile<0>::ROWS), g.n[2]);
        hipLaunchKernelGGL((k_sweep<R, 0, MODE>), grid, dim3(256), 0, c->stream, a);
    } else if (d == 1) {
        dim3 grid(blocks_for(g.n[0], SweepTile<1>::LANES), blocks_for(nif, SweepTile<1>::ROWS), g.n[2]);
        hipLaunchKernelGGL((k_sweep<R, 1, MODE>), grid, dim3(256), 0, c->stream, a);
    } else {
        dim3 grid(blocks_for(g.n[0], SweepTile<2>::LANES), blocks_for(nif, SweepTile<2>::ROWS), g.n[1]);
        hipLaunchKernelGGL((k_sweep<R, 2, MODE>), grid, dim3(256), 0, c->stream, a);
    }
}

template <typename R>
SweepArgs<R> sweep_args(howmhd_gpu_ctx* c, int d, const double* qsrc) {
    SweepArgs<R> a{};
    a.q = qsrc;
    a.dF = (R*)c->dF[d];
    // x: fsy, fsz; y: gsz, gsx; z: hsx, hsy
    a.ct_a = (R*)c->ct[2 * d];
    a.ct_b = (R*)c->ct[2 * d + 1];
    if (d == 1) { a.ct_a = (R*)c->ct[3]; a.ct_b = (R*)c->ct[2]; }
    a.g = c->g;
    a.ph = c->ph;
    a.mlp = c->mlp;
    a.stats = (c->ind_mode == 2) ? c->istats : nullptr;
    return a;
}

template <typename R>
int launch_sweeps(howmhd_gpu_ctx* c, const double* qsrc) {
    for (int d = 0; d < 3; ++d) {
        if (c->g.n[d] <= 1) continue;
        const SweepArgs<R> a = sweep_args<R>(c, d, qsrc);
        TimedLaunch tl(c, HOWMHD_K_SWEEP_X + d);
        if (c->ind_mode) launch_sweep<R, SW_ROUTED>(c, d, a);
        else launch_sweep<R, SW_PLAIN>(c, d, a);
        HCHECK(c, hipGetLastError());
    }
    return HOWMHD_OK;
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

template <typename R>
int launch_stage(howmhd_gpu_ctx* c, int iw, double dt) {
    const RK k = rk_coeff(iw);
    StageCoef s;
    s.k1 = k.k1;
    s.k3 = k.k3;
    s.c4 = (iw == 5) ? k.k4 / k.k1 : 0.0;
    s.c5 = (iw == 5) ? k.k5 / k.k1 : 0.0;
    s.dtdx = k.k2 * dt / c->p.dx;   // src/ssprk.cpp:1302-1304
    s.dtdy = k.k2 * dt / c->p.dy;
    s.dtdz = k.k2 * dt / c->p.dz;
    s.stage5 = (iw == 5);

    const double* q0 = c->qbuf[c->qid[0]];
    const double* qp = c->qbuf[c->qid[iw - 1]];
    double* qo = c->qbuf[c->qid[iw]];
    const double* b0 = c->fbuf[c->qid[0]];
    const double* bp = c->fbuf[c->qid[iw - 1]];
    double* bo = c->fbuf[c->qid[iw]];

    int rc = launch_sweeps<R>(c, qp);
    if (rc) return rc;

    const unsigned nb = blocks_for(c->ncell, 256);
    if (c->ctf.active) {
        EmfArgs<R> e;
        e.fsy = (const R*)c->ct[0]; e.fsz = (const R*)c->ct[1];
        e.gsx = (const R*)c->ct[2]; e.gsz = (const R*)c->ct[3];
        e.hsx = (const R*)c->ct[4]; e.hsy = (const R*)c->ct[5];
        e.Ox2 = (R*)c->O2[0]; e.Oy2 = (R*)c->O2[1]; e.Oz2 = (R*)c->O2[2];
        e.g = c->g;
        e.f = c->ctf;
        {
            TimedLaunch tl(c, HOWMHD_K_EMF2);
            hipLaunchKernelGGL((k_emf2<R>), dim3(nb), dim3(256), 0, c->stream, e);
        }
        HCHECK(c, hipGetLastError());
        SmoothArgs<R> m;
        m.Ox2 = (const R*)c->O2[0]; m.Oy2 = (const R*)c->O2[1]; m.Oz2 = (const R*)c->O2[2];
        m.Ox = (R*)c->O[0]; m.Oy = (R*)c->O[1]; m.Oz = (R*)c->O[2];
        m.g = c->g;
        m.f = c->ctf;
        {
            TimedLaunch tl(c, HOWMHD_K_EMF);
            hipLaunchKernelGGL((k_emf<R>), dim3(nb), dim3(256), 0, c->stream, m);
        }
        HCHECK(c, hipGetLastError());
    }
    FaceArgs<R> fa;
    fa.Ox = (const R*)c->O[0]; fa.Oy = (const R*)c->O[1]; fa.Oz = (const R*)c->O[2];
    fa.b0 = b0;
    fa.b2 = c->fbuf[c->qid[2]];
    fa.b3 = c->fbuf[c->qid[3]];
    fa.bp = bp;
    fa.bo = bo;
    fa.g = c->g;
    fa.f = c->ctf;
    fa.s = s;
    {
        TimedLaunch tl(c, HOWMHD_K_FACES);
        hipLaunchKernelGGL((k_faces<R>), dim3(nb), dim3(256), 0, c->stream, fa);
    }
    HCHECK(c, hipGetLastError());

    CombineArgs<R> ca;
    ca.q0 = q0;
    ca.q2 = c->qbuf[c->qid[2]];
    ca.q3 = c->qbuf[c->qid[3]];
    ca.qp = qp;
    ca.qo = qo;
    ca.bo = bo;
    ca.dFx = (const R*)c->dF[0];
    ca.dFy = (const R*)c->dF[1];
    ca.dFz = (const R*)c->dF[2];
    ca.g = c->g;
    ca.f = c->ctf;
    ca.s = s;
    ca.ph = c->ph;
    {
        TimedLaunch tl(c, HOWMHD_K_COMBINE);
        hipLaunchKernelGGL((k_combine<R>), dim3(nb), dim3(256), 0, c->stream, ca);
    }
    HCHECK(c, hipGetLastError());
    return HOWMHD_OK;
}

// reference layout (m, i, j, k) with ghosts <-> device layout (m, z, y, x) active
void ref_to_dev(const howmhd_gpu_ctx* c, const double* ref, double* dev, int nvar) {
    const int nx = c->g.n[0], ny = c->g.n[1], nz = c->g.n[2];
    const size_t n6y = (size_t)ny + 6, n6z = (size_t)nz + 6, n6 = ((size_t)nx + 6) * n6y * n6z;
    for (int m = 0; m < nvar; ++m)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double* src = ref + (size_t)m * n6 + (((size_t)x + 3) * n6y + (size_t)y + 3) * n6z + 3;
                double* dst = dev + (size_t)m * c->ncell + (size_t)y * nx + x;
                for (int z = 0; z < nz; ++z) dst[(size_t)z * nx * ny] = src[z];
            }
}

int host_map(int a, int n, int per) {
    if (n == 1) return 0;
    if (per) return ((a % n) + n) % n;
    return a < 0 ? 0 : (a >= n ? n - 1 : a);
}

void dev_to_ref(const howmhd_gpu_ctx* c, const double* dev, double* ref, int nvar) {
    const int nx = c->g.n[0], ny = c->g.n[1], nz = c->g.n[2];
    const size_t n6x = (size_t)nx + 6, n6y = (size_t)ny + 6, n6z = (size_t)nz + 6, n6 = n6x * n6y * n6z;
    std::vector<int> mx(n6x), my(n6y), mz(n6z);
    for (size_t i = 0; i < n6x; ++i) mx[i] = host_map((int)i - 3, nx, c->g.per[0]);
    for (size_t j = 0; j < n6y; ++j) my[j] = host_map((int)j - 3, ny, c->g.per[1]);
    for (size_t k = 0; k < n6z; ++k) mz[k] = host_map((int)k - 3, nz, c->g.per[2]);
    for (int m = 0; m < nvar; ++m)
        for (size_t i = 0; i < n6x; ++i)
            for (size_t j = 0; j < n6y; ++j) {
                double* dst = ref + (size_t)m * n6 + (i * n6y + j) * n6z;
                const double* src = dev + (size_t)m * c->ncell + (size_t)my[j] * nx + mx[i];
                for (size_t k = 0; k < n6z; ++k) dst[k] = src[(size_t)mz[k] * nx * ny];
            }
}

}  // namespace

extern "C" {

int howmhd_gpu_create(const howmhd_params* p, howmhd_gpu_ctx** out) {
    if (!p || !out) {
        g_create_error = "howmhd_gpu_create: null argument";
        return HOWMHD_EINVAL;
    }
    *out = nullptr;
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int d = 0; d < 3; ++d) {
        if (n[d] < 1 || n[d] == 2) {
            g_create_error = "howmhd_gpu_create: each axis needs n == 1 or n >= 3 active cells";
            return HOWMHD_EINVAL;
        }
        if (p->bc[d] != HOWMHD_BC_PERIODIC && p->bc[d] != HOWMHD_BC_OPEN) {
            g_create_error = "howmhd_gpu_create: unknown boundary condition code";
            return HOWMHD_EINVAL;
        }
    }
    if (p->precision != HOWMHD_PREC_FP64 && p->precision != HOWMHD_PREC_FP32) {
        g_create_error = "howmhd_gpu_create: unknown precision";
        return HOWMHD_EINVAL;
    }
    if (p->indicator != HOWMHD_INDICATOR_OFF && p->indicator != HOWMHD_INDICATOR_MLP) {
        g_create_error = "howmhd_gpu_create: unknown indicator";
        return HOWMHD_EINVAL;
    }
    if (!(p->dx > 0 && p->dy > 0 && p->dz > 0) || !(p->gam > 1.0)) {
        g_create_error = "howmhd_gpu_create: dx, dy, dz must be > 0 and gam > 1";
        return HOWMHD_EINVAL;
    }
    howmhd_gpu_ctx* c = new howmhd_gpu_ctx();
    c->p = *p;
    for (int d = 0; d < 3; ++d) {
        c->g.n[d] = n[d];
        c->g.per[d] = (p->bc[d] == HOWMHD_BC_PERIODIC) ? 1 : 0;
    }
    c->ncell = (size_t)n[0] * n[1] * n[2];
    c->g.ncell = (long long)c->ncell;
    c->g.sx = 1; c->g.sy = n[0]; c->g.sz = n[0] * n[1];
    c->ph.gam = p->gam; c->ph.rhomin = p->rhomin; c->ph.pgmin = p->pgmin;
    c->ctf.use_x = n[0] > 1; c->ctf.use_y = n[1] > 1; c->ctf.use_z = n[2] > 1;
    c->ctf.use_Ox = c->ctf.use_x && c->ctf.use_y;
    c->ctf.use_Oy = c->ctf.use_y && c->ctf.use_z;
    c->ctf.use_Oz = c->ctf.use_x && c->ctf.use_z;
    c->ctf.active = c->ctf.use_Ox || c->ctf.use_Oy || c->ctf.use_Oz;
    c->rsize = (p->precision == HOWMHD_PREC_FP32) ? sizeof(float) : sizeof(double);

    auto cleanup_fail = [&](int code, const std::string& msg) {
        g_create_error = msg;
        howmhd_gpu_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return cleanup_fail(HOWMHD_EHIP, "hipGetDevice failed (no GPU?)");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(HOWMHD_EHIP, "hipStreamCreate failed");
    auto alloc = [&](void** ptr, size_t bytes) -> bool {
        if (hipMalloc(ptr, bytes) != hipSuccess) return false;
        return hipMemset(*ptr, 0, bytes) == hipSuccess;
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
    for (int t = 0; t < 3 && ok; ++t) {
        ok = ok && alloc(&c->O2[t], c->ncell * c->rsize);
        ok = ok && alloc(&c->O[t], c->ncell * c->rsize);
    }
    ok = ok && alloc((void**)&c->dmax, 4 * sizeof(unsigned long long));
    ok = ok && alloc((void**)&c->mlp, sizeof(MlpDev));
    ok = ok && alloc((void**)&c->istats, 64 * 4 * sizeof(unsigned long long));
    if (!ok) return cleanup_fail(HOWMHD_ENOMEM, "howmhd_gpu_create: device allocation failed");
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
    if (c->mlp) (void)hipFree(c->mlp);
    if (c->istats) (void)hipFree(c->istats);
    drain_timing(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HOWMHD_OK;
}

int howmhd_gpu_upload_q0(howmhd_gpu_ctx* c, const double* q0) {
    if (!c || !q0) return HOWMHD_EINVAL;
    std::vector<double> h(8 * c->ncell);
    ref_to_dev(c, q0, h.data(), 8);
    HCHECK(c, hipMemcpyAsync(c->qbuf[c->qid[0]], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->faces_ready = false;
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
    std::vector<double> h(3 * c->ncell);
    ref_to_dev(c, b, h.data(), 3);
    HCHECK(c, hipMemcpyAsync(c->fbuf[c->qid[0]], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->faces_ready = true;
    return HOWMHD_OK;
}

int howmhd_gpu_tstep(howmhd_gpu_ctx* c, double* dt) {
    if (!c || !dt) return HOWMHD_EINVAL;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->dmax, 3);
    HCHECK(c, hipGetLastError());
    const unsigned nb = std::min<unsigned>(blocks_for(c->ncell, 256), 2048u);
    {
        TimedLaunch tl(c, HOWMHD_K_CFL);
        hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, c->stream, (const double*)c->qbuf[c->qid[0]], c->g, c->ph,
                           c->dmax);
    }
    HCHECK(c, hipGetLastError());
    unsigned long long h[3];
    HCHECK(c, hipMemcpyAsync(h, c->dmax, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    double vmax[3];
    for (int d = 0; d < 3; ++d) std::memcpy(&vmax[d], &h[d], sizeof(double));
    // src/tstep.cpp:123-135
    const double tiny = 1.0e-300;
    double invdt = 0.0;
    if (c->p.nx > 1) invdt += std::fmax(vmax[0], tiny) / c->p.dx;
    if (c->p.ny > 1) invdt += std::fmax(vmax[1], tiny) / c->p.dy;
    if (c->p.nz > 1) invdt += std::fmax(vmax[2], tiny) / c->p.dz;
    if (invdt <= 0.0) invdt = std::fmax(vmax[0], tiny) / c->p.dx;
    *dt = c->p.cour / invdt;
    return HOWMHD_OK;
}

int howmhd_gpu_ssprk(howmhd_gpu_ctx* c, double dt) {
    if (!c) return HOWMHD_EINVAL;
    if (!c->faces_ready) return fail(c, HOWMHD_ESTATE, "howmhd_gpu_ssprk: call init_faces or upload_faces first");
    if (c->p.indicator == HOWMHD_INDICATOR_MLP && !c->ind_mode)
        return fail(c, HOWMHD_ESTATE, "howmhd_gpu_ssprk: indicator mlp needs weights (howmhd_gpu_set_indicator)");
    for (int iw = 1; iw <= 5; ++iw) {
        const int rc = (c->rsize == sizeof(float)) ? launch_stage<float>(c, iw, dt) : launch_stage<double>(c, iw, dt);
        if (rc) return rc;
    }
    // q(0) <- q(5), faces likewise: exchange the buffers of stages 0 and 5
    std::swap(c->qid[0], c->qid[1]);
    c->qid[5] = c->qid[1];
    return HOWMHD_OK;
}

int howmhd_gpu_download_q0(howmhd_gpu_ctx* c, double* q0) {
    if (!c || !q0) return HOWMHD_EINVAL;
    std::vector<double> h(8 * c->ncell);
    HCHECK(c, hipMemcpyAsync(h.data(), c->qbuf[c->qid[0]], h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    dev_to_ref(c, h.data(), q0, 8);
    return HOWMHD_OK;
}

int howmhd_gpu_download_faces(howmhd_gpu_ctx* c, double* b) {
    if (!c || !b) return HOWMHD_EINVAL;
    std::vector<double> h(3 * c->ncell);
    HCHECK(c, hipMemcpyAsync(h.data(), c->fbuf[c->qid[0]], h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    dev_to_ref(c, h.data(), b, 3);
    return HOWMHD_OK;
}

int howmhd_gpu_diag(howmhd_gpu_ctx* c, howmhd_diag* d) {
    if (!c || !d) return HOWMHD_EINVAL;
    std::memset(d, 0, sizeof(*d));
    const size_t nc = c->ncell;
    std::vector<double> q(8 * nc), b(3 * nc);
    HCHECK(c, hipMemcpyAsync(q.data(), c->qbuf[c->qid[0]], q.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipMemcpyAsync(b.data(), c->fbuf[c->qid[0]], b.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    const int nx = c->g.n[0], ny = c->g.n[1], nz = c->g.n[2];
    const double gam = c->p.gam;
    d->min_rho = 1e300;
    d->min_p = 1e300;
    for (int m = 0; m < 8; ++m) {
        double s = 0.0;
        for (size_t i = 0; i < nc; ++i) s += q[m * nc + i];
        d->sum_cons[m] = s;
    }
    auto M = [&](int a, int n, int per) { return host_map(a, n, per); };
    const double aa = 9.0 / 8.0, bb = -1.0 / 24.0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = ((size_t)z * ny + y) * nx + x;
                const double rho = q[i];
                const double vx = q[nc + i] / rho, vy = q[2 * nc + i] / rho, vz = q[3 * nc + i] / rho;
                const double BB2 = q[4 * nc + i] * q[4 * nc + i] + q[5 * nc + i] * q[5 * nc + i] + q[6 * nc + i] * q[6 * nc + i];
                const double p = (gam - 1.0) * (q[7 * nc + i] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * BB2);
                d->min_rho = std::min(d->min_rho, rho);
                d->min_p = std::min(d->min_p, p);
                // CT face divergence: D4 of src/fluxct.cpp:848 along each axis
                double dv = 0.0;
                const int P0 = c->g.per[0], P1 = c->g.per[1], P2 = c->g.per[2];
                if (nx > 1) {
                    auto bx = [&](int xx) { return b[((size_t)z * ny + y) * nx + M(xx, nx, P0)]; };
                    dv += (aa * (bx(x) - bx(x - 1)) + bb * (bx(x + 1) - bx(x - 2))) / c->p.dx;
                }
                if (ny > 1) {
                    auto by = [&](int yy) { return b[nc + ((size_t)z * ny + M(yy, ny, P1)) * nx + x]; };
                    dv += (aa * (by(y) - by(y - 1)) + bb * (by(y + 1) - by(y - 2))) / c->p.dy;
                }
                if (nz > 1) {
                    auto bz = [&](int zz) { return b[2 * nc + ((size_t)M(zz, nz, P2) * ny + y) * nx + x]; };
                    dv += (aa * (bz(z) - bz(z - 1)) + bb * (bz(z + 1) - bz(z - 2))) / c->p.dz;
                }
                d->max_divb_ct = std::max(d->max_divb_ct, std::fabs(dv));
                // 2nd-order cell-centred diagnostic, src/output.cpp:314-347
                double dc = 0.0;
                if (nx > 1) {
                    const int xm = M(x - 1, nx, P0), xp = M(x + 1, nx, P0);
                    const double den = (xm == x || xp == x) ? c->p.dx : 2.0 * c->p.dx;
                    dc += (q[4 * nc + ((size_t)z * ny + y) * nx + xp] - q[4 * nc + ((size_t)z * ny + y) * nx + xm]) / den;
                }
                if (ny > 1) {
                    const int ym = M(y - 1, ny, P1), yp = M(y + 1, ny, P1);
                    const double den = (ym == y || yp == y) ? c->p.dy : 2.0 * c->p.dy;
                    dc += (q[5 * nc + ((size_t)z * ny + yp) * nx + x] - q[5 * nc + ((size_t)z * ny + ym) * nx + x]) / den;
                }
                if (nz > 1) {
                    const int zm = M(z - 1, nz, P2), zp = M(z + 1, nz, P2);
                    const double den = (zm == z || zp == z) ? c->p.dz : 2.0 * c->p.dz;
                    dc += (q[6 * nc + ((size_t)zp * ny + y) * nx + x] - q[6 * nc + ((size_t)zm * ny + y) * nx + x]) / den;
                }
                d->max_divb_cell = std::max(d->max_divb_cell, std::fabs(dc));
            }
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
    if (c->rsize != sizeof(double))
        return fail(c, HOWMHD_EINVAL, "howmhd_gpu_collect_samples: labels come from the FP64 solver (precision fp64)");
    // the sweep writes this direction's dF and CT terms too; the next stage
    // recomputes them from its own source
    SweepArgs<double> a = sweep_args<double>(c, dir, c->qbuf[c->qid[0]]);
    a.feat = X;
    a.label = y;
    a.stats = nullptr;
    launch_sweep<double, SW_COLLECT>(c, dir, a);
    HCHECK(c, hipGetLastError());
    HCHECK(c, hipStreamSynchronize(c->stream));
    *count = need;
    return HOWMHD_OK;
}

int howmhd_gpu_set_indicator(howmhd_gpu_ctx* c, const float* params, float threshold, int enable) {
    if (!c) return HOWMHD_EINVAL;
    if (enable < 0 || enable > 2) return fail(c, HOWMHD_EINVAL, "howmhd_gpu_set_indicator: enable is 0, 1 or 2");
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
        for (int k = 0; k < MLP_H; ++k)
            for (int nn = 0; nn < MLP_H; ++nn) {
                h.W1t[k][nn] = k < MLP_IN ? bf16_bits(params[MLP_OFF_W1 + nn * MLP_IN + k]) : 0;
                h.W2t[k][nn] = bf16_bits(params[MLP_OFF_W2 + nn * MLP_H + k]);
            }
        for (int j = 0; j < MLP_H; ++j) {
            h.b1[j] = params[MLP_OFF_B1 + j];
            h.b2[j] = params[MLP_OFF_B2 + j];
            h.w3[j] = bf16_round(params[MLP_OFF_W3 + j]);
        }
        h.b3 = params[MLP_OFF_B3];
        h.logit_threshold = std::log(threshold / (1.f - threshold));
        HCHECK(c, hipMemcpy(c->mlp, &h, sizeof(h), hipMemcpyHostToDevice));
        HCHECK(c, hipMemset(c->istats, 0, 64 * 4 * sizeof(unsigned long long)));
    }
    c->ind_mode = enable;
    return HOWMHD_OK;
}

int howmhd_gpu_indicator_stats(howmhd_gpu_ctx* c, int64_t* flagged, int64_t* total, int64_t* waves_flagged,
                               int64_t* waves_total) {
    if (!c || !flagged || !total || !waves_flagged || !waves_total) return HOWMHD_EINVAL;
    unsigned long long h[64 * 4];
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(h, c->istats, sizeof(h), hipMemcpyDeviceToHost));
    HCHECK(c, hipMemset(c->istats, 0, sizeof(h)));
    int64_t s[4] = {0, 0, 0, 0};
    for (int i = 0; i < 64; ++i)
        for (int k = 0; k < 4; ++k) s[k] += (int64_t)h[4 * i + k];
    *flagged = s[0];
    *total = s[1];
    *waves_flagged = s[2];
    *waves_total = s[3];
    return HOWMHD_OK;
}

const char* howmhd_gpu_build_kind(void) {
#if HOWMHD_EXACT
    return "fp-exact";
#else
    return "fp-fast";
#endif
}

}  // extern "C"
