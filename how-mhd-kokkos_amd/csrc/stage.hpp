// Launch sequence of one SSPRK(5,4) stage (src/ssprk.cpp:1277-1329 with
// fluxct, src/fluxct.cpp:753-899).  Included by each translation unit of the
// library; the kernels it launches carry EX = HOWMHD_EXACT of that unit.
#pragma once

#include "context.hpp"

#include <algorithm>

namespace howmhd {
// internal linkage: each translation unit keeps its own copy (the units
// differ in HOWMHD_EXACT and contraction)
namespace {

constexpr int EX = HOWMHD_EXACT;

struct RK {
    double k1, k2, k3, k4, k5;
};

// SSPRK(5,4) coefficients, src/ssprk.cpp:66-94
inline RK rk_coeff(int iw) {
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

template <typename R, int MODE, int DM>
void launch_sweep_dm(howmhd_gpu_ctx* c, int d, const SweepArgs<R>& a, dim3 grid) {
    if (d == 0) hipLaunchKernelGGL((k_sweep<R, 0, MODE, EX, DM>), grid, dim3(256), 0, c->stream, a);
    else if (d == 1) hipLaunchKernelGGL((k_sweep<R, 1, MODE, EX, DM>), grid, dim3(256), 0, c->stream, a);
    else hipLaunchKernelGGL((k_sweep<R, 2, MODE, EX, DM>), grid, dim3(256), 0, c->stream, a);
}

// Blocks of the IEEE-division recompute launch of the reference-rounding form.
constexpr unsigned kRecomputeBlocks = 64;

// Tiles a y/z-sweep block marches through (sweep_march): as many as keep at
// least kMarchMinBlocks blocks in the launch, at most HOWMHD_MARCH_STEPS.
#ifndef HOWMHD_MARCH_STEPS
#define HOWMHD_MARCH_STEPS 8
#endif
constexpr unsigned kMarchMinBlocks = 2048;
inline int march_steps(const dim3& tiles) {
    const unsigned long long n = (unsigned long long)tiles.x * tiles.y * tiles.z;
    unsigned long long s = n / kMarchMinBlocks;
    if (s > HOWMHD_MARCH_STEPS) s = HOWMHD_MARCH_STEPS;
    if (s > tiles.y) s = tiles.y;
    return s < 1 ? 1 : (int)s;
}

// The sweep along d.  Reference rounding: the sweep with unscaled WENO
// divisions, then the IEEE-division launch for the tiles it queued (it exits
// at once when there are none; no counting there).
template <typename R, int MODE>
void launch_sweep(howmhd_gpu_ctx* c, int d, const SweepArgs<R>& a0) {
    SweepArgs<R> a = a0;
    dim3 grid = sweep_grid(c->g, d);
    a.march = 1;
    if (d != 0 && sweep_marches<R, MODE>()) {
        a.march = c->march_force > 0 ? (int)std::min<unsigned>(c->march_force, grid.y) : march_steps(grid);
        grid.y = (grid.y + a.march - 1) / a.march;
    }
    if constexpr (EX) {
        launch_sweep_dm<R, MODE, DIV_UNSCALED>(c, d, a, grid);
        SweepArgs<R> b = a;
        b.stats = nullptr;
        launch_sweep_dm<R, MODE, DIV_IEEE>(c, d, b, dim3(kRecomputeBlocks));
    } else {
        launch_sweep_dm<R, MODE, DIV_IEEE>(c, d, a, grid);
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
    a.gc = c->gc;
    a.ct_active = c->ctf.active;
    a.mlp = c->mlp;
    a.stats = c->route_now ? c->iwave : nullptr;
    a.fix = c->fixflag;
    a.fixlist = c->fixlist;
    a.fixctl = c->fixctl;
    return a;
}

inline StageCoef stage_coef(const howmhd_gpu_ctx* c, int iw, double dt) {
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
    return s;
}

// The stage combination of stage iw (its inputs: q(0), q(iw-1), stage 2 and
// 3 at stage 5, the faces and interface fluxes of stage iw; its output q(iw)).
template <typename R>
CombineArgs<R> combine_args(howmhd_gpu_ctx* c, int iw, double dt) {
    CombineArgs<R> ca{};
    ca.q0 = c->qbuf[c->qid[0]];
    ca.q2 = c->qbuf[c->qid[2]];
    ca.q3 = c->qbuf[c->qid[3]];
    ca.qp = c->qbuf[c->qid[iw - 1]];
    ca.qo = c->qbuf[c->qid[iw]];
    ca.bo = c->fbuf[c->qid[iw]];
    ca.dFx = (const R*)c->dF[0];
    ca.dFy = (const R*)c->dF[1];
    ca.dFz = (const R*)c->dF[2];
    ca.g = c->g;
    ca.f = c->ctf;
    ca.s = stage_coef(c, iw, dt);
    ca.ph = c->ph;
    ca.cfl = (iw == 5) ? c->cflpart : nullptr;
    ca.floors = (iw == 5) ? c->floorpart : nullptr;
    return ca;
}

template <typename R>
int launch_sweeps(howmhd_gpu_ctx* c, const double* qsrc) {
    for (int d = 0; d < 3; ++d) {
        if (c->g.n[d] <= 1) continue;
        const SweepArgs<R> a = sweep_args<R>(c, d, qsrc);
        TimedLaunch tl(c, HOWMHD_K_SWEEP_X + d);
        if (c->route_now) launch_sweep<R, SW_ROUTED>(c, d, a);
        else launch_sweep<R, SW_PLAIN>(c, d, a);
        HCHECK(c, hipGetLastError());
    }
    return HOWMHD_OK;
}

template <typename R>
int launch_stage(howmhd_gpu_ctx* c, int iw, double dt) {
    const StageCoef s = stage_coef(c, iw, dt);
    const double* qp = c->qbuf[c->qid[iw - 1]];
    const double* b0 = c->fbuf[c->qid[0]];
    const double* bp = c->fbuf[c->qid[iw - 1]];
    double* bo = c->fbuf[c->qid[iw]];

    int rc = launch_sweeps<R>(c, qp);
    if (rc) return rc;

    const unsigned nb = blocks_for(c->ncell, 256);
    if (c->ctf.use_x && c->ctf.use_y && c->ctf.use_z) {
        EmfMarchArgs<R> e;
        e.fsy = (const R*)c->ct[0]; e.fsz = (const R*)c->ct[1];
        e.gsx = (const R*)c->ct[2]; e.gsz = (const R*)c->ct[3];
        e.hsx = (const R*)c->ct[4]; e.hsy = (const R*)c->ct[5];
        e.Ox = (R*)c->O[0]; e.Oy = (R*)c->O[1]; e.Oz = (R*)c->O[2];
        e.g = c->g;
        const dim3 grid(blocks_for(c->g.n[0], EMF_OX), blocks_for(c->g.n[1], EMF_OY), blocks_for(c->g.n[2], EMF_CZ));
        {
            TimedLaunch tl(c, HOWMHD_K_EMF);
            hipLaunchKernelGGL((k_emf_march<R, EX>), grid, dim3(EMF_THREADS), 0, c->stream, e);
        }
        HCHECK(c, hipGetLastError());
    } else if (c->ctf.active) {
        EmfArgs<R> e;
        e.fsy = (const R*)c->ct[0]; e.fsz = (const R*)c->ct[1];
        e.gsx = (const R*)c->ct[2]; e.gsz = (const R*)c->ct[3];
        e.hsx = (const R*)c->ct[4]; e.hsy = (const R*)c->ct[5];
        e.Ox2 = (R*)c->O2[0]; e.Oy2 = (R*)c->O2[1]; e.Oz2 = (R*)c->O2[2];
        e.g = c->g;
        e.f = c->ctf;
        {
            TimedLaunch tl(c, HOWMHD_K_EMF2);
            hipLaunchKernelGGL((k_emf2<R, EX>), dim3(nb), dim3(256), 0, c->stream, e);
        }
        HCHECK(c, hipGetLastError());
        SmoothArgs<R> m;
        m.Ox2 = (const R*)c->O2[0]; m.Oy2 = (const R*)c->O2[1]; m.Oz2 = (const R*)c->O2[2];
        m.Ox = (R*)c->O[0]; m.Oy = (R*)c->O[1]; m.Oz = (R*)c->O[2];
        m.g = c->g;
        m.f = c->ctf;
        {
            TimedLaunch tl(c, HOWMHD_K_EMF);
            hipLaunchKernelGGL((k_emf<R, EX>), dim3(nb), dim3(256), 0, c->stream, m);
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
        if (c->ctf.use_x && c->ctf.use_y && c->ctf.use_z) {
            const dim3 fgrid(blocks_for(c->g.n[0], FACE_TX), blocks_for(c->g.n[1], FACE_TY),
                             blocks_for(c->g.n[2], FACE_CZ));
            hipLaunchKernelGGL((k_faces_march<R, EX>), fgrid, dim3(256), 0, c->stream, fa);
        } else {
            const dim3 fgrid(blocks_for(c->g.n[0], FACE_TX), blocks_for(c->g.n[1], FACE_TY), c->g.n[2]);
            hipLaunchKernelGGL((k_faces<R, EX>), fgrid, dim3(256), 0, c->stream, fa);
        }
    }
    HCHECK(c, hipGetLastError());

    {
        const CombineArgs<R> ca = combine_args<R>(c, iw, dt);
        TimedLaunch tl(c, HOWMHD_K_COMBINE);
        hipLaunchKernelGGL((k_combine<R, EX>), dim3(nb), dim3(256), 0, c->stream, ca);
    }
    HCHECK(c, hipGetLastError());
    return HOWMHD_OK;
}

}  // namespace
}  // namespace howmhd
