// HIP kernels of the HOW-MHD time step for gfx950.
//
// Device layout: only active cells are stored, one array per variable,
// x fastest: index (z*ny + y)*nx + x.  A neighbour outside the active range is
// read through the boundary map (periodic wrap or open clamp), which yields
// exactly the ghost values src/bound.cpp and the serial fills of
// src/fluxct.cpp would have written (axes with n >= 3; a degenerate n = 1 axis
// is never stencilled by the time step).
//
// Per SSPRK stage iw the kernels are:
//   k_sweep<X|Y|Z>  fused primit + eigensystem + characteristic LLF-WENO5:
//                   interface fluxes dF (7 per interface) and CT flux terms
//                   (src/ssprk.cpp:458-997, 1000-1272)
//   k_emf_march     corner EMF build and smoothing (src/fluxct.cpp:809-840)
//                   in one z-marching pass on 3-D grids; k_emf2, k_emf as two
//                   passes on 1-D and 2-D grids
//   k_faces         face-B RK combination + 4th-order staggered curl
//                   (src/fluxct.cpp:848-877, 111-135); k_faces_march, the
//                   same marching through z on 3-D grids
//   k_combine       stage combination k1*q_init + k3*q_prev + dX + dY + dZ,
//                   face->centre B (src/fluxct.cpp:882-897) and the floors of
//                   prot (src/prot.cpp:10-51)
// Each kernel carries EX = HOWMHD_EXACT of the translation unit that
// instantiates it (the plain FP64 unit: reference rounding; the FMA unit:
// fp64fast and fp32), so the two units never share a kernel symbol.  The
// per-step kernels (CFL, face initialisation) are in tstep_kernels.hpp.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "mhd_math.hpp"
#include "mlp.hpp"

namespace howmhd {

// ------------------------------------------------------------------ geometry
struct Geom {
    int n[3];      // active cells per axis
    int per[3];    // 1 periodic, 0 open
    long long ncell;
};

// Active index a in [-3, n+2] -> stored index, following src/bound.cpp.
__device__ __forceinline__ int bmap(int a, int n, int per) {
    if (n == 1) return 0;
    if (per) return a < 0 ? a + n : (a >= n ? a - n : a);
    return a < 0 ? 0 : (a >= n ? n - 1 : a);
}

// Coordinates of stored index i, and the stored index of (x, y, z).
__device__ __forceinline__ void cell_of(const Geom& g, long long i, int& x, int& y, int& z) {
    x = (int)(i % g.n[0]);
    const long long r = i / g.n[0];
    y = (int)(r % g.n[1]);
    z = (int)(r / g.n[1]);
}
__device__ __forceinline__ long long lin(const Geom& g, int x, int y, int z) {
    return ((long long)z * g.n[1] + y) * g.n[0] + x;
}

// Streaming stores and loads.  The arrays a kernel writes here (interface
// fluxes, CT terms, EMFs, faces, q) are read again only by a later kernel,
// after more than the L2 and the 256 MB Infinity Cache have passed through,
// so they are written non-temporally and leave those caches to the stencil
// re-reads; the stage combination reads q(0) and q(iw-1) once the same way.
// Measured at 256^3 (FP64): face update 0.368 -> 0.345 ms, combination
// 1.029 -> 1.004 ms, step 36.8 -> 36.7 ms.
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }

// Sweep permutation of src/ssprk.cpp:381-403, 1035-1065, 1126-1156: sweep slot m
// holds q variable sweep_perm(dir, m); order (rho, Mn, Mt1, Mt2, Bn, Bt1, Bt2,
// E): x {0,1,2,3,4,5,6,7}, y {0,2,3,1,5,6,4,7}, z {0,3,1,2,6,4,5,7}.  A
// constant expression, so every use folds to an immediate offset.
__host__ __device__ constexpr int sweep_perm(int dir, int m) {
    return (m == 0 || m == 7) ? m : (m < 4 ? 1 + (m - 1 + dir) % 3 : 4 + (m - 4 + dir) % 3);
}
static_assert(sweep_perm(1, 1) == 2 && sweep_perm(1, 3) == 1 && sweep_perm(1, 6) == 4 &&
              sweep_perm(2, 1) == 3 && sweep_perm(2, 2) == 1 && sweep_perm(2, 4) == 6 &&
              sweep_perm(0, 5) == 5 && sweep_perm(2, 7) == 7, "sweep permutation");

struct Phys {
    double gam, rhomin, pgmin;
};

// ------------------------------------------------------------- sweep tiling
// X sweep: 64 interfaces along x per row, 4 rows (y) per block.
// Y/Z sweep: HOWMHD_YZ_LANES x-columns, 256/HOWMHD_YZ_LANES interfaces along the
// sweep axis per block.
template <int DIR> struct SweepTile;
template <> struct SweepTile<0> {
    static constexpr int LANES = 64, ROWS = 4, CELLS = (64 + 5) * 4;
};
#ifndef HOWMHD_YZ_LANES
#define HOWMHD_YZ_LANES 32
#endif
template <> struct SweepTile<1> {
    static constexpr int LANES = HOWMHD_YZ_LANES, ROWS = 256 / HOWMHD_YZ_LANES, CELLS = LANES * (ROWS + 5);
};
template <> struct SweepTile<2> {
    static constexpr int LANES = HOWMHD_YZ_LANES, ROWS = 256 / HOWMHD_YZ_LANES, CELLS = LANES * (ROWS + 5);
};

// The minimum number of wavefronts per SIMD the FP64 sweeps are compiled for
// (caps their registers at 512 / that number).
#ifndef HOWMHD_SWEEP_WAVES
#define HOWMHD_SWEEP_WAVES 2
#endif
// Stencil cells whose LDS reads the scheduler may overlap; a scheduling
// barrier after every such batch keeps it from hoisting all 42 reads of a
// pass ahead of the arithmetic (which spills registers).  One cell per batch:
// the FP64 sweeps then need 182-212 registers and the FP32 sweeps 93-113,
// none spilled; two cells per batch spill in every form.
#ifndef HOWMHD_PROJ_BATCH
#define HOWMHD_PROJ_BATCH 1
#endif

__host__ __device__ inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Launch grid of the sweep along axis d (256-thread blocks; SweepTile).
__host__ __device__ inline dim3 sweep_grid(const Geom& g, int d) {
    const int nif = g.per[d] ? g.n[d] : g.n[d] + 1;
    if (d == 0) return dim3(blocks_for(nif, SweepTile<0>::LANES), blocks_for(g.n[1], SweepTile<0>::ROWS), g.n[2]);
    if (d == 1) return dim3(blocks_for(g.n[0], SweepTile<1>::LANES), blocks_for(nif, SweepTile<1>::ROWS), g.n[2]);
    return dim3(blocks_for(g.n[0], SweepTile<2>::LANES), blocks_for(nif, SweepTile<2>::ROWS), g.n[1]);
}

// Sweep variants: plain characteristic WENO5; the same plus indicator
// samples (features and the WENO-weight label); routed by the learned
// indicator (flagged interfaces: characteristic WENO5, others: linear
// upwind-5 on the components).
enum SweepMode { SW_PLAIN = 0, SW_COLLECT = 1, SW_ROUTED = 2 };

template <typename R>
struct SweepArgs {
    const double* q;    // source stage, 8 arrays of ncell
    R* dF;              // 7 x n_if planes (see dF_index)
    // constrained transport active: components 4, 5 of dF (the
    // transverse-B fluxes) feed only the CT terms, not k_combine, and are
    // not stored
    int ct_active;
    R* ct_a;            // CT term "bsy": fsy (x), gsz (y), hsx (z)
    R* ct_b;            // CT term "bsz": fsz (x), gsx (y), hsy (z)
    Geom g;
    Phys ph;
    GasConst gc;
    // SW_COLLECT: 18 features and the label per interface (sample_index), and
    // (when logit is set) the in-sweep network's logit
    float* feat;
    float* label;
    float* logit;
    // SW_ROUTED: device weights; stats (null: not counted): one slot per
    // wavefront of the launch grid (block-linear index x 4 + wavefront), to
    // which the wavefront adds (flagged, total, 1 if any flagged, 1): plain
    // read-modify-writes (every slot has one writer per launch) that
    // k_fold_wave_stats sums and clears after the step
    const MlpDev* mlp;
    uint4* stats;
    // reference-rounding form: tiles whose WENO weight operands left the
    // range of the unscaled division, for the IEEE-division launch that
    // follows: a flag per tile (block-linear index, so a tile is queued
    // once), the queue, and (queue length, blocks done) of that launch
    unsigned* fix;
    unsigned* fixlist;
    unsigned* fixctl;
    // y/z sweeps: consecutive tiles along the sweep axis per block
    // (sweep_march; 1: single tiles)
    int march;
};

// dF layout per direction: component-major, interface index t = 0..n along the
// sweep axis (t is the left face of active cell t):
//   x: ((c*nz + z)*ny + y)*(nx+1) + t
//   y: ((c*nz + z)*(ny+1) + t)*nx + x
//   z: ((c*(nz+1) + t)*ny + y)*nx + x
template <int DIR>
__device__ __forceinline__ long long dF_index(const Geom& g, int c, int t, int x, int y, int z) {
    if constexpr (DIR == 0) return (((long long)c * g.n[2] + z) * g.n[1] + y) * (g.n[0] + 1) + t;
    else if constexpr (DIR == 1) return (((long long)c * g.n[2] + z) * (g.n[1] + 1) + t) * g.n[0] + x;
    else return (((long long)c * (g.n[2] + 1) + t) * g.n[1] + y) * g.n[0] + x;
}

// Dense index of a sampled interface: the lines of the sweep, each with its
// computed interfaces t = tbeg..n (tbeg = 1 periodic, 0 open) in order.
template <int DIR>
__device__ __forceinline__ long long sample_index(const Geom& g, int t, int x, int y, int z) {
    const int tbeg = g.per[DIR] ? 1 : 0;
    const long long nint = g.n[DIR] + 1 - tbeg, u = t - tbeg;
    if constexpr (DIR == 0) return ((long long)z * g.n[1] + y) * nint + u;
    else if constexpr (DIR == 1) return ((long long)z * nint + u) * g.n[0] + x;
    else return (u * g.n[1] + y) * g.n[0] + x;
}

template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

// (U_j, F_j) of one cell side by side, so one 128-bit LDS read serves both.
template <typename R>
struct alignas(2 * sizeof(R)) UF {
    R u, f;
};

// Per-cell fields besides U and F, in LDS.
enum OtherField { O_VN = 0, O_LF, O_LA, O_LS, O_VT1, O_VT2, O_BN, O_P, O_H, O_COUNT };

template <typename R>
__device__ __forceinline__ void opq(R& x) { asm volatile("" : "+v"(x)); }

// Makes x opaque to the optimiser and, through the input operand, ordered
// after `after` (no instruction emitted).
template <typename R>
__device__ __forceinline__ void opq_after(R& x, R after) { asm volatile("" : "+v"(x) : "v"(after)); }

// The scalars the R columns use, made opaque after the wave flux exists: the
// column is then built after the WENO step instead of being kept live across it.
template <typename R>
__device__ __forceinline__ void opaque_after(Eig<R>& o, R a) {
    opq_after(o.vx, a); opq_after(o.vy, a); opq_after(o.vz, a); opq_after(o.vv2, a);
    opq_after(o.cs, a); opq_after(o.cs2, a); opq_after(o.lf, a); opq_after(o.ls, a);
    opq_after(o.af, a); opq_after(o.as, a); opq_after(o.bty, a); opq_after(o.btz, a);
    opq_after(o.sgnBx, a); opq_after(o.invsqrtrho, a);
}

// Makes every scalar of e opaque to the optimiser (no instruction emitted).
template <typename R>
__device__ __forceinline__ void opaque(Eig<R>& o) {
    opq(o.vx); opq(o.vy); opq(o.vz); opq(o.By); opq(o.Bz); opq(o.vv2); opq(o.cs); opq(o.cs2);
    opq(o.lf); opq(o.ls); opq(o.af); opq(o.as); opq(o.bty); opq(o.btz); opq(o.sgnBx);
    opq(o.sqrtrho); opq(o.invsqrtrho); opq(o.nrm_half); opq(o.nrm_ent);
}

// Characteristic LLF-WENO5 flux of the interface of phase-2 thread (row,
// lane) from the sweep tile in LDS (src/ssprk.cpp:825-946): interface
// eigensystem, the 7 wave maxima, then the waves in pairs, added to dFo in
// wave order.  The WENO weights divide in mode DM (DivMode); `bad` collects
// the lanes whose operands left the range of the unscaled division.
template <typename R, int DIR, int MODE, int DM>
__device__ __forceinline__ void weno_interface(const UF<R> (*luf)[SweepTile<DIR>::CELLS],
                                               const R (*lo)[SweepTile<DIR>::CELLS], int row, int lane, R gam,
                                               const GasConst& gc, R (&dFo)[7], R& mdev, unsigned long long& bad) {
    using T = SweepTile<DIR>;
    (void)mdev;
    auto slot = [&](int k) -> int {
        if constexpr (DIR == 0) return row * (T::LANES + 5) + lane + k;
        else return (row + k) * T::LANES + lane;
    };
    // interface state from cells k = 2, 3: (rho, vn, vt1, vt2, Bn, Bt1, Bt2, p, H)
    R u0[9], u1[9];
    {
        const int s0 = slot(2), s1 = slot(3);
        u0[0] = luf[0][s0].u; u0[1] = lo[O_VN][s0]; u0[2] = lo[O_VT1][s0]; u0[3] = lo[O_VT2][s0];
        u0[4] = lo[O_BN][s0]; u0[5] = luf[4][s0].u; u0[6] = luf[5][s0].u; u0[7] = lo[O_P][s0]; u0[8] = lo[O_H][s0];
        u1[0] = luf[0][s1].u; u1[1] = lo[O_VN][s1]; u1[2] = lo[O_VT1][s1]; u1[3] = lo[O_VT2][s1];
        u1[4] = lo[O_BN][s1]; u1[5] = luf[4][s1].u; u1[6] = luf[5][s1].u; u1[7] = lo[O_P][s1]; u1[8] = lo[O_H][s1];
    }
    Eig<R> e;
    bool sqbad = false;
    eig_scalars<R, DM>(gam, u0, u1, e, &sqbad);
    // the unscaled mode normalises the L rows with div_pos (needs 1/cs2 > 0)
    // and takes unscaled square roots (operands 0 or >= 2^-767)
    if constexpr (DM == DIV_UNSCALED) bad |= __ballot(!(e.nrm_half > R(0)) || sqbad);

    // The waves are projected in four passes over the stencil in LDS,
    // (0,6), (1,5), (2,4) and (3), and dF still gets them added in order 0..6
    // (the products of waves 4, 5, 6 with their R columns wait for wave 3).
    // A pass shares work between its two rows of L (src/ssprk.cpp:662-716)
    // and its two columns of R (729-787):
    //  * rows 0 and 6, and rows 2 and 4, have equal entries 4..6, so those
    //    three products of a stencil cell serve both rows;
    //  * rows 1 and 5 have equal entries 0, 2, 3, opposite entries 4, 5 and
    //    zeros at 1 and 6: both sums share the partial sum over entries 0, 2,
    //    3, row 5 subtracts row 1's last two products, and the zero products
    //    are left out.
    // A product left out is +-0 and a product reused with its sign changed
    // is exact, so a projected value can differ from the reference's only in
    // the sign of a zero.  The wave's flux then differs at most in the sign
    // of a zero, and a zero product added to dF leaves it unchanged (dF
    // starts at +0; a sum of doubles is -0 only when both addends are).  For
    // the same reason the structural zeros of R (columns 1 and 5 in rows 0,
    // 1; column 3 in rows 4, 5) add nothing and are skipped.  The product
    // R[r][m] * F_s is rounded once whenever it is formed, so forming the
    // products of waves 4-6 early and adding them after wave 3 gives dF the
    // same bits.
    auto add_col = [&](auto mc, const R* Rc, R Fs) {
        constexpr int M = decltype(mc)::value;
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            if ((M == 1 || M == 5) && r < 2) continue;
            if (M == 3 && (r == 4 || r == 5)) continue;
            dFo[r] += Rc[r] * Fs;
        }
    };
    // products R[r][m] * F_s of waves 4, 5, 6, waiting for wave 3
    R P4[7], P5[7], P6[7];
    auto add_prod = [&](auto mc, const R* Pm) {
        constexpr int M = decltype(mc)::value;
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            if (M == 5 && r < 2) continue;
            dFo[r] += Pm[r];
        }
    };
    sfor<0, 4>([&](auto pc) {
        constexpr int P = decltype(pc)::value;
        constexpr int ma = P;                // 0, 1, 2, 3
        constexpr int mb = 6 - P;            // 6, 5, 4 (3: none)
        constexpr bool two = (P < 3);
        // the stencil is read from LDS again for every pass rather than kept
        // live in registers
        asm volatile("" ::: "memory");
        // the scalars become opaque here, so the compiler builds this pass's
        // rows and columns here instead of all seven up front (98 live values)
        opaque(e);
        // largest |eigenvalue| of the pass's waves over the 6 cells
        // (src/ssprk.cpp:851-854): v_n -+ c_f, v_n -+ c_a, v_n -+ c_s, v_n
        R amxa = R(0), amxb = R(0);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int s = slot(k);
            const R vn = lo[O_VN][s];
            if constexpr (P == 3) {
                amxa = rmax(amxa, rabs(vn));
            } else {
                const R c = lo[P == 0 ? O_LF : P == 1 ? O_LA : O_LS][s];
                amxa = rmax(amxa, rabs(vn - c));
                amxb = rmax(amxb, rabs(vn + c));
            }
        }
        R La[7], Lb[7];
        l_row<ma, R, DM>(gam, e, La, gc);
        if constexpr (P == 0 || P == 2) l_row<mb, R, DM>(gam, e, Lb, gc);
        R Fa[6], qa[6], Fb[6], qb[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int s = slot(k);
            R U[7], F[7];
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const UF<R> v = luf[j][s];
                U[j] = v.u;
                F[j] = v.f;
            }
            if constexpr (P == 1) {
                const R pf = La[0] * F[0] + La[2] * F[2] + La[3] * F[3];
                const R pu = La[0] * U[0] + La[2] * U[2] + La[3] * U[3];
                const R f4 = La[4] * F[4], f5 = La[5] * F[5];
                const R u4 = La[4] * U[4], u5 = La[5] * U[5];
                Fa[k] = pf + f4 + f5;
                qa[k] = pu + u4 + u5;
                Fb[k] = pf - f4 - f5;
                qb[k] = pu - u4 - u5;
            } else if constexpr (two) {
                const R f4 = La[4] * F[4], f5 = La[5] * F[5], f6 = La[6] * F[6];
                const R u4 = La[4] * U[4], u5 = La[5] * U[5], u6 = La[6] * U[6];
                Fa[k] = La[0] * F[0] + La[1] * F[1] + La[2] * F[2] + La[3] * F[3] + f4 + f5 + f6;
                qa[k] = La[0] * U[0] + La[1] * U[1] + La[2] * U[2] + La[3] * U[3] + u4 + u5 + u6;
                Fb[k] = Lb[0] * F[0] + Lb[1] * F[1] + Lb[2] * F[2] + Lb[3] * F[3] + f4 + f5 + f6;
                qb[k] = Lb[0] * U[0] + Lb[1] * U[1] + Lb[2] * U[2] + Lb[3] * U[3] + u4 + u5 + u6;
            } else {
                Fa[k] = La[0] * F[0] + La[1] * F[1] + La[2] * F[2] + La[3] * F[3] + La[4] * F[4] + La[5] * F[5] + La[6] * F[6];
                qa[k] = La[0] * U[0] + La[1] * U[1] + La[2] * U[2] + La[3] * U[3] + La[4] * U[4] + La[5] * U[5] + La[6] * U[6];
            }
            if ((k + 1) % HOWMHD_PROJ_BATCH == 0) asm volatile("" ::: "memory");
        }
        R Fsa, Fsb = R(0);
        if constexpr (MODE == SW_COLLECT) {
            Fsa = weno_wave_dev<R, DM>(Fa, qa, amxa, mdev, bad);
            if constexpr (two) Fsb = weno_wave_dev<R, DM>(Fb, qb, amxb, mdev, bad);
        } else {
            Fsa = weno_wave<R, DM>(Fa, qa, amxa, bad);
            if constexpr (two) Fsb = weno_wave<R, DM>(Fb, qb, amxb, bad);
        }
        opaque_after(e, Fsa);
        if constexpr (two) opaque_after(e, Fsb);
        R Rc[7];
        r_col<ma>(gam, e, Rc, gc);
        add_col(std::integral_constant<int, ma>{}, Rc, Fsa);
        if constexpr (two) {
            // column mb shares most of its products with column ma (built
            // together, so the compiler forms them once); its products with
            // the wave flux are added to dF after wave 3's
            R Rb[7];
            r_col<mb>(gam, e, Rb, gc);
            R* Pb = (P == 0) ? P6 : (P == 1) ? P5 : P4;
#pragma unroll
            for (int r = 0; r < 7; ++r) Pb[r] = Rb[r] * Fsb;
        } else {
            add_prod(std::integral_constant<int, 4>{}, P4);
            add_prod(std::integral_constant<int, 5>{}, P5);
            add_prod(std::integral_constant<int, 6>{}, P6);
        }
    });
}

// Wavefronts per SIMD the sweep is compiled for: FP32 routed sweeps 4 (at
// most 128 registers; four blocks' tiles fill the 160 KB of LDS), the others
// HOWMHD_SWEEP_WAVES.
#ifndef HOWMHD_ROUTED_WAVES
#define HOWMHD_ROUTED_WAVES 4
#endif
template <typename R, int MODE>
constexpr int sweep_waves() { return (sizeof(R) == 4 && MODE != SW_COLLECT) ? HOWMHD_ROUTED_WAVES : HOWMHD_SWEEP_WAVES; }

// y/z sweeps that march through consecutive tiles (sweep_march): the FP64
// ones but the collecting sweep.  The FP32 sweep at 4 waves per SIMD and the
// collecting sweep have no registers for the march's state and would spill.
template <typename R, int MODE>
constexpr bool sweep_marches() { return sizeof(R) == 8 && MODE != SW_COLLECT; }

// Stored index of slot s of the LDS tile whose first interface is t0 (the
// tile's cells run from t0 - 3 to t0 + ROWS + 1 along the sweep axis, x
// sweep: t0 - 3 to t0 + LANES + 1), through the boundary map.
template <int DIR>
__device__ __forceinline__ long long tile_cell_index(const Geom& g, int s, int t0, int x0, int y0, int z0) {
    using T = SweepTile<DIR>;
    const int nd = g.n[DIR];
    int cx, cy, cz, along;
    if constexpr (DIR == 0) {
        const int row = s / (T::LANES + 5), c = s % (T::LANES + 5);
        along = t0 - 3 + c;
        cy = y0 + row; cz = z0; cx = 0;
        if (cy >= g.n[1]) cy = g.n[1] - 1;
    } else {
        const int row = s / T::LANES, lane = s % T::LANES;
        along = t0 - 3 + row;
        cx = x0 + lane;
        if (cx >= g.n[0]) cx = g.n[0] - 1;
        if constexpr (DIR == 1) { cz = z0; cy = 0; }
        else { cy = y0; cz = 0; }
    }
    if (along > nd + 2) along = nd + 2;
    const int a = bmap(along, nd, g.per[DIR]);
    if constexpr (DIR == 0) cx = a;
    else if constexpr (DIR == 1) cy = a;
    else cz = a;
    return ((long long)cz * g.n[1] + cy) * g.n[0] + cx;
}

// Cell values of one cell (its 8 conserved values in sweep order) into LDS
// slot s.
template <typename R, int DIR, int MODE, int DM>
__device__ __forceinline__ void store_cell(const SweepArgs<R>& A, UF<R> (*luf)[SweepTile<DIR>::CELLS],
                                           R (*lo)[SweepTile<DIR>::CELLS], float* lbmag, int s, const double* q,
                                           bool& cell_bad) {
    R out[C_NFIELD];
    cell_values<R, DM>(q, A.ph.gam, A.ph.rhomin, A.ph.pgmin, out, &cell_bad);
#pragma unroll
    for (int j = 0; j < 7; ++j) luf[j][s] = UF<R>{out[C_RHO + j], out[C_F0 + j]};
#pragma unroll
    for (int f = 0; f < O_COUNT; ++f) lo[f][s] = out[C_VN + f];
    if constexpr (MODE != SW_PLAIN)
        lbmag[s] = sqrtf((float)(out[C_BN] * out[C_BN] + out[C_BT1] * out[C_BT1] + out[C_BT2] * out[C_BT2]));
}

// Phase 2 of a tile whose cell values are in LDS: one interface per thread
// (t0, x0, y0, z0: the tile origin; blk: its index in the single-tile grid).
template <typename R, int DIR, int MODE, int EX, int DM>
__device__ __forceinline__ void sweep_interfaces(const SweepArgs<R>& A, const UF<R> (*luf)[SweepTile<DIR>::CELLS],
                                                 const R (*lo)[SweepTile<DIR>::CELLS], const float* lbmag, int t0,
                                                 int x0, int y0, int z0, long long blk) {
    using T = SweepTile<DIR>;
    const Geom& g = A.g;
    const int nd = g.n[DIR];
    const int tid = threadIdx.x;
    int lane, row, t, x, y, z;
    bool valid;
    if constexpr (DIR == 0) {
        lane = tid % T::LANES; row = tid / T::LANES;
        t = t0 + lane; x = 0; y = y0 + row; z = z0;
        valid = (t <= nd && y < g.n[1]);
    } else {
        lane = tid % T::LANES; row = tid / T::LANES;
        t = t0 + row; x = x0 + lane;
        if constexpr (DIR == 1) { y = 0; z = z0; }
        else { y = y0; z = 0; }
        valid = (t <= nd && x < g.n[0]);
    }
    auto slot = [&](int k) -> int {
        if constexpr (DIR == 0) return row * (T::LANES + 5) + lane + k;
        else return (row + k) * T::LANES + lane;
    };
    // indicator features: the 6-cell stencils of rho, p and |B|
    auto features = [&](float* f) {
        float rr[6], pp[6], bb[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int s = valid ? slot(k) : 0;
            rr[k] = (float)luf[0][s].u;
            pp[k] = (float)lo[O_P][s];
            bb[k] = lbmag[s];
        }
        mlp_features(rr, pp, bb, f);
    };
    bool flagged = true;
    if constexpr (MODE == SW_COLLECT) {
        if (A.logit) {   // uniform over the launch
            float f[MLP_IN];
            features(f);
            const float zl = mlp_forward_wave(f, *A.mlp);
            if (valid) A.logit[sample_index<DIR>(g, t, x, y, z)] = zl;
        }
    }
    if constexpr (MODE == SW_ROUTED) {
        float f[MLP_IN];
        features(f);
        const float zl = mlp_forward_wave(f, *A.mlp);
        flagged = valid && (zl > A.mlp->logit_threshold);
        if (A.stats) {
            const unsigned long long fb = __ballot(flagged), vb = __ballot(valid);
            if ((tid & 63) == 0 && vb) {
                uint4* st = A.stats + 4 * blk + (tid >> 6);
                uint4 v = *st;
                v.x += (unsigned)__popcll(fb);
                v.y += (unsigned)__popcll(vb);
                v.z += fb ? 1u : 0u;
                v.w += 1u;
                *st = v;
            }
        }
    }
    if (!valid) return;

    R dF[7] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
    R maxdev = R(0);
    if (MODE != SW_ROUTED || flagged) {
        unsigned long long bad = 0ull;
        weno_interface<R, DIR, MODE, DM>(luf, lo, row, lane, R(A.ph.gam), A.gc, dF, maxdev, bad);
        // a WENO weight operand beyond the unscaled-division range: the tile
        // goes (once) on the list of the IEEE-division launch that follows
        if constexpr (EX && DM == DIV_UNSCALED) {
            if (bad && atomicExch(&A.fix[blk], 1u) == 0u) A.fixlist[atomicAdd(&A.fixctl[0], 1u)] = (unsigned)blk;
        }
    } else {
        // linear upwind-5 with one LLF speed: the largest |v_n| + c_f over the
        // stencil bounds every wave's speed
        R alpha = R(0);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int s = slot(k);
            alpha = rmax(alpha, rabs(lo[O_VN][s]) + lo[O_LF][s]);
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            R U[6], F[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const UF<R> v = luf[j][slot(k)];
                U[k] = v.u;
                F[k] = v.f;
            }
            dF[j] = upwind5<R>(F, U, alpha);
        }
    }

#pragma unroll
    for (int c = 0; c < 7; ++c)
        if (!(A.ct_active && (c == 4 || c == 5))) st_nt(&A.dF[dF_index<DIR>(g, c, t, x, y, z)], dF[c]);

    // CT flux terms, src/ssprk.cpp:967-994: (-0 a + 9 b + 9 c - 0 d)/18 over
    // cells k = 1..4.  The products with the zero weights are +-0 and are
    // left out: the sum can then differ only in the sign of a zero, which
    // adding it to dF[4] or dF[5] (never -0) does not show.
    if (t >= 1) {
        R bnv1[2], bnv2[2];
#pragma unroll
        for (int k = 2; k <= 3; ++k) {
            const int s = slot(k);
            const R bn = lo[O_BN][s];
            bnv1[k - 2] = bn * lo[O_VT1][s];
            bnv2[k - 2] = bn * lo[O_VT2][s];
        }
        const R ftmp1 = div_c<R>(R(9) * bnv1[0] + R(9) * bnv1[1], R(18), R(1.0 / 18.0));
        const R ftmp2 = div_c<R>(R(9) * bnv2[0] + R(9) * bnv2[1], R(18), R(1.0 / 18.0));
        long long ci;
        if constexpr (DIR == 0) ci = ((long long)z * g.n[1] + y) * g.n[0] + (t - 1);
        else if constexpr (DIR == 1) ci = ((long long)z * g.n[1] + (t - 1)) * g.n[0] + x;
        else ci = ((long long)(t - 1) * g.n[1] + y) * g.n[0] + x;
        st_nt(&A.ct_a[ci], dF[4] + ftmp1);
        st_nt(&A.ct_b[ci], dF[5] + ftmp2);
    }
    if constexpr (MODE == SW_COLLECT) {
        if (!A.feat) return;
        float f[MLP_IN];
        features(f);
        const long long si = sample_index<DIR>(g, t, x, y, z);
#pragma unroll
        for (int j = 0; j < MLP_IN; ++j) A.feat[si * MLP_IN + j] = f[j];
        A.label[si] = (maxdev > R(0.05)) ? 1.f : 0.f;
    }
}

// One tile (bx, by, bz) of the sweep grid (sweep_grid), block-linear index
// blk.  DM (DivMode) of the WENO weights: the FMA forms and the
// reference-rounding recompute launch divide with IEEE divisions; the
// reference-rounding sweep proper with div_recip/div_with, queueing its tile
// for the recompute launch where an operand leaves their range.
template <typename R, int DIR, int MODE, int EX, int DM>
__device__ __forceinline__ void sweep_tile(const SweepArgs<R>& A, int bx, int by, int bz, long long blk) {
    using T = SweepTile<DIR>;
    __shared__ UF<R> luf[7][T::CELLS];
    __shared__ R lo[O_COUNT][T::CELLS];
    // |B| of each cell for the indicator features (routed and sample sweeps)
    __shared__ float lbmag[MODE != SW_PLAIN ? T::CELLS : 1];
    const Geom& g = A.g;
    const int tbeg = g.per[DIR] ? 1 : 0;  // periodic: interface 0 is the image of n
    const int tid = threadIdx.x;

    // block origin
    int t0, x0 = 0, y0 = 0, z0 = 0;
    if constexpr (DIR == 0) {
        t0 = tbeg + bx * T::LANES;
        y0 = by * T::ROWS;
        z0 = bz;
    } else {
        x0 = bx * T::LANES;
        t0 = tbeg + by * T::ROWS;
        if constexpr (DIR == 1) z0 = bz;
        else y0 = bz;
    }

    // ---- phase 1: cell values of the tile (stencil halo included) into LDS.
    // A thread's cells are tid, tid + 256, ...; all their loads are issued
    // before the first conversion, so a block waits on memory once.
    auto cell_index = [&](int s) -> long long { return tile_cell_index<DIR>(g, s, t0, x0, y0, z0); };
    constexpr int NIT = (T::CELLS + 255) / 256;
    double qs[NIT][8];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int s = tid + 256 * it;
        if (s < T::CELLS) {
            const long long idx = cell_index(s);
#pragma unroll
            for (int m = 0; m < 8; ++m) qs[it][m] = A.q[(long long)sweep_perm(DIR, m) * g.ncell + idx];
        }
    }
    bool cell_bad = false;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int s = tid + 256 * it;
        if (s < T::CELLS) store_cell<R, DIR, MODE, DM>(A, luf, lo, lbmag, s, qs[it], cell_bad);
    }
    // a cell without positive shared reciprocals: the tile goes (once) on the
    // list of the IEEE-division launch that follows
    if constexpr (EX && DM == DIV_UNSCALED) {
        if (__ballot(cell_bad) && atomicExch(&A.fix[blk], 1u) == 0u)
            A.fixlist[atomicAdd(&A.fixctl[0], 1u)] = (unsigned)blk;
    }
    __syncthreads();

    sweep_interfaces<R, DIR, MODE, EX, DM>(A, luf, lo, lbmag, t0, x0, y0, z0, blk);
}

// The y or z sweep of up to `nsteps` consecutive tiles along the sweep axis
// (tiles by0, by0 + 1, ... of the single-tile grid) by one block.
// Consecutive tiles share 5 of their 13 cell rows: each step after the first
// moves rows 8..12 of the previous tile to rows 0..4 of the LDS tile and
// converts only its 8 new rows, one cell per thread.  Every value is the one
// sweep_tile computes for the same tile; queueing and statistics use the
// tile's single-tile index.
template <typename R, int DIR, int MODE, int EX, int DM>
__device__ __forceinline__ void sweep_march(const SweepArgs<R>& A, int bx, int by0, int bz, int nsteps) {
    static_assert(DIR != 0, "the x sweep takes single tiles");
    using T = SweepTile<DIR>;
    constexpr int HALO = 5 * T::LANES;        // slots of rows 0..4
    constexpr int NEW = T::CELLS - HALO;      // slots of rows 5..12
    constexpr int KEEP = T::ROWS * T::LANES;  // first slot of rows 8..12
    static_assert(NEW == 256 && KEEP + HALO == T::CELLS, "one new cell per thread");
    __shared__ UF<R> luf[7][T::CELLS];
    __shared__ R lo[O_COUNT][T::CELLS];
    __shared__ float lbmag[MODE != SW_PLAIN ? T::CELLS : 1];
    const Geom& g = A.g;
    const int tbeg = g.per[DIR] ? 1 : 0;
    const int tid = threadIdx.x;
    const dim3 tg = sweep_grid(g, DIR);  // the single-tile grid
    const int x0 = bx * T::LANES;
    int y0 = 0, z0 = 0;
    if constexpr (DIR == 1) z0 = bz;
    else y0 = bz;
    static_assert(sweep_marches<R, MODE>(), "a sweep with registers for the march");
    // Loads of the next step's cell issued before the interfaces of this
    // step: only where the 16 registers they hold leave the interfaces
    // without spills (the plain sweep of reference rounding); elsewhere each
    // step loads its cell at its start.
    constexpr bool EARLY = EX && MODE == SW_PLAIN;
    double qn[8];            // this thread's new cell of the next step
    bool bad_prev = false;   // this thread converted a cell out of range in the previous step
    auto load_new = [&](int t0) {  // the new cell of the tile starting at row t0
        const long long idx = tile_cell_index<DIR>(g, HALO + tid, t0, x0, y0, z0);
#pragma unroll
        for (int m = 0; m < 8; ++m) qn[m] = A.q[(long long)sweep_perm(DIR, m) * g.ncell + idx];
    };
    // queueing of a tile (see sweep_tile) and the early loads
    auto finish_cells = [&](bool cell_bad, int by, int t0, long long blk) {
        if constexpr (EX && DM == DIV_UNSCALED) {
            if (__ballot(cell_bad || bad_prev) && atomicExch(&A.fix[blk], 1u) == 0u)
                A.fixlist[atomicAdd(&A.fixctl[0], 1u)] = (unsigned)blk;
        }
        bad_prev = cell_bad;
        if constexpr (EARLY) {
            if (by + 1 < by0 + nsteps && by + 1 < (int)tg.y) load_new(t0 + T::ROWS);
        }
    };
    if (by0 >= (int)tg.y) return;  // uniform over the block
    // first tile: every cell, as sweep_tile
    {
        const int t0 = tbeg + by0 * T::ROWS;
        const long long blk = bx + (long long)tg.x * (by0 + (long long)tg.y * bz);
        bool cell_bad = false;
        constexpr int NIT = (T::CELLS + 255) / 256;
        double qs[NIT][8];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int s = tid + 256 * it;
            if (s < T::CELLS) {
                const long long idx = tile_cell_index<DIR>(g, s, t0, x0, y0, z0);
#pragma unroll
                for (int m = 0; m < 8; ++m) qs[it][m] = A.q[(long long)sweep_perm(DIR, m) * g.ncell + idx];
            }
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int s = tid + 256 * it;
            if (s < T::CELLS) store_cell<R, DIR, MODE, DM>(A, luf, lo, lbmag, s, qs[it], cell_bad);
        }
        finish_cells(cell_bad, by0, t0, blk);
        __syncthreads();  // the tile's cells are in LDS
        sweep_interfaces<R, DIR, MODE, EX, DM>(A, luf, lo, lbmag, t0, x0, y0, z0, blk);
    }
    for (int by = by0 + 1; by < by0 + nsteps && by < (int)tg.y; ++by) {  // uniform over the block
        const int t0 = tbeg + by * T::ROWS;
        const long long blk = bx + (long long)tg.x * (by + (long long)tg.y * bz);
        if constexpr (!EARLY) load_new(t0);
        __syncthreads();  // every interface of the previous tile has read it
        // rows 8..12 of the previous tile are rows 0..4 of this one (source
        // and destination slots are disjoint)
        if (tid < HALO) {
            const int src = KEEP + tid;
            UF<R> u[7];
            R o[O_COUNT];
#pragma unroll
            for (int j = 0; j < 7; ++j) u[j] = luf[j][src];
#pragma unroll
            for (int f = 0; f < O_COUNT; ++f) o[f] = lo[f][src];
#pragma unroll
            for (int j = 0; j < 7; ++j) luf[j][tid] = u[j];
#pragma unroll
            for (int f = 0; f < O_COUNT; ++f) lo[f][tid] = o[f];
            if constexpr (MODE != SW_PLAIN) lbmag[tid] = lbmag[src];
        }
        __syncthreads();  // rows 8..12 are read before the new rows overwrite them
        bool cell_bad = false;
        store_cell<R, DIR, MODE, DM>(A, luf, lo, lbmag, HALO + tid, qn, cell_bad);
        finish_cells(cell_bad, by, t0, blk);
        __syncthreads();  // the tile's cells are in LDS
        sweep_interfaces<R, DIR, MODE, EX, DM>(A, luf, lo, lbmag, t0, x0, y0, z0, blk);
    }
}

// ---------------------------------------------------------------- CT kernels
struct CtFlags {
    int use_x, use_y, use_z, use_Ox, use_Oy, use_Oz, active;
};

// Tile of an x-sweep workgroup.  Workgroups are dispatched round-robin over
// the 8 XCDs, and each XCD has its own L2.  In launch order the x-tiles of one
// row would land on different XCDs, and each would fetch the cache lines of
// the 3-cell stencil halo that it shares with its neighbours (1.5x the row's
// bytes).  Here workgroups L and L + 8 (the same XCD, dispatched one after
// the other) take neighbouring x-tiles of one row; rows beyond the last full
// group of 8 keep the launch order.  A bijection on the grid.
__device__ __forceinline__ void xcd_row_tiles(int& bx, int& by, int& bz) {
    const unsigned gx = gridDim.x, gy = gridDim.y, nrow = gy * gridDim.z;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    bx = (int)blockIdx.x; by = (int)blockIdx.y; bz = (int)blockIdx.z;
    if (gx == 1 || L >= (nrow / 8) * 8 * gx) return;
    const unsigned a = L / 8, row = (a / gx) * 8 + L % 8;
    bx = (int)(a % gx); by = (int)(row % gy); bz = (int)(row / gy);
}

// The sweep along DIR.  Reference rounding, DM = DIV_IEEE: the recompute
// launch (a few blocks) takes the tiles queued by the sweep proper
// (A.fixlist, A.fixctl[0] of them), clears their flags, and the last block
// to finish empties the queue; with an empty queue every block exits at once.
template <typename R, int DIR, int MODE, int EX, int DM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sweep_waves<R, MODE>()))) void k_sweep(
    SweepArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    static_assert(EX || DM == DIV_IEEE, "the FMA forms divide with IEEE divisions");
    if constexpr (EX && DM == DIV_IEEE) {
        const unsigned n = A.fixctl[0];
        const dim3 gr = sweep_grid(A.g, DIR);
        for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
            const unsigned b = A.fixlist[i];
            __syncthreads();  // the previous tile's LDS is free
            sweep_tile<R, DIR, MODE, EX, DM>(A, (int)(b % gr.x), (int)((b / gr.x) % gr.y), (int)(b / (gr.x * gr.y)), b);
            if (threadIdx.x == 0) A.fix[b] = 0u;
        }
        __syncthreads();  // every thread of the block has read the count
        if (threadIdx.x == 0 && atomicAdd(&A.fixctl[1], 1u) == gridDim.x - 1) {
            A.fixctl[0] = 0u;  // every block has read the count
            A.fixctl[1] = 0u;
        }
    } else {
        int bx = (int)blockIdx.x, by = (int)blockIdx.y, bz = (int)blockIdx.z;
        if constexpr (DIR != 0 && sweep_marches<R, MODE>()) {
            sweep_march<R, DIR, MODE, EX, DM>(A, bx, by * A.march, bz, A.march);
        } else {
            if constexpr (DIR == 0) xcd_row_tiles(bx, by, bz);
            const long long blk = bx + (long long)gridDim.x * (by + (long long)gridDim.y * bz);
            sweep_tile<R, DIR, MODE, EX, DM>(A, bx, by, bz, blk);
        }
    }
}

template <typename R>
struct EmfArgs {
    const R *fsy, *fsz, *gsx, *gsz, *hsx, *hsy;
    R *Ox2, *Oy2, *Oz2;
    Geom g;
    CtFlags f;
};


// 4-point interpolation of src/fluxct.cpp:815-820: (-a + 9b + 9c - d)/16.
template <typename R>
__device__ __forceinline__ R i4(R a, R b, R c, R d) {
    return div_c<R>(-a + R(9) * b + R(9) * c - d, R(16), R(1.0 / 16.0));
}

// fluxct_build_O2, src/fluxct.cpp:809-821, on active cells
template <typename R, int EX>
__global__ __launch_bounds__(256) void k_emf2(EmfArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const Geom& g = A.g;
    if (i >= g.ncell) return;
    int x, y, z;
    cell_of(g, i, x, y, z);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    if (A.f.use_Ox) {
        const int xm = bmap(x - 1, nx, g.per[0]), xp = bmap(x + 1, nx, g.per[0]), xq = bmap(x + 2, nx, g.per[0]);
        const int ym = bmap(y - 1, ny, g.per[1]), yp = bmap(y + 1, ny, g.per[1]), yq = bmap(y + 2, ny, g.per[1]);
        A.Ox2[i] = i4<R>(A.gsx[lin(g, xm, y, z)], A.gsx[i], A.gsx[lin(g, xp, y, z)], A.gsx[lin(g, xq, y, z)]) -
                   i4<R>(A.fsy[lin(g, x, ym, z)], A.fsy[i], A.fsy[lin(g, x, yp, z)], A.fsy[lin(g, x, yq, z)]);
    }
    if (A.f.use_Oy) {
        const int ym = bmap(y - 1, ny, g.per[1]), yp = bmap(y + 1, ny, g.per[1]), yq = bmap(y + 2, ny, g.per[1]);
        const int zm = bmap(z - 1, nz, g.per[2]), zp = bmap(z + 1, nz, g.per[2]), zq = bmap(z + 2, nz, g.per[2]);
        A.Oy2[i] = i4<R>(A.hsy[lin(g, x, ym, z)], A.hsy[i], A.hsy[lin(g, x, yp, z)], A.hsy[lin(g, x, yq, z)]) -
                   i4<R>(A.gsz[lin(g, x, y, zm)], A.gsz[i], A.gsz[lin(g, x, y, zp)], A.gsz[lin(g, x, y, zq)]);
    }
    if (A.f.use_Oz) {
        const int zm = bmap(z - 1, nz, g.per[2]), zp = bmap(z + 1, nz, g.per[2]), zq = bmap(z + 2, nz, g.per[2]);
        const int xm = bmap(x - 1, nx, g.per[0]), xp = bmap(x + 1, nx, g.per[0]), xq = bmap(x + 2, nx, g.per[0]);
        A.Oz2[i] = i4<R>(A.fsz[lin(g, x, y, zm)], A.fsz[i], A.fsz[lin(g, x, y, zp)], A.fsz[lin(g, x, y, zq)]) -
                   i4<R>(A.hsx[lin(g, xm, y, z)], A.hsx[i], A.hsx[lin(g, xp, y, z)], A.hsx[lin(g, xq, y, z)]);
    }
}

// (a - 2c + b)/24 of src/fluxct.cpp:834-839
template <typename R>
__device__ __forceinline__ R lap24(R m, R c, R p) {
    return div_c<R>(m - R(2) * c + p, R(24), R(1.0 / 24.0));
}

template <typename R>
struct SmoothArgs {
    const R *Ox2, *Oy2, *Oz2;
    R *Ox, *Oy, *Oz;
    Geom g;
    CtFlags f;
};

// fluxct_smooth_O, src/fluxct.cpp:828-840
template <typename R, int EX>
__global__ __launch_bounds__(256) void k_emf(SmoothArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const Geom& g = A.g;
    if (i >= g.ncell) return;
    int x, y, z;
    cell_of(g, i, x, y, z);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    if (A.f.use_Ox) {
        const R o = A.Ox2[i];
        A.Ox[i] = o + lap24<R>(A.Ox2[lin(g, bmap(x - 1, nx, g.per[0]), y, z)], o, A.Ox2[lin(g, bmap(x + 1, nx, g.per[0]), y, z)]) +
                  lap24<R>(A.Ox2[lin(g, x, bmap(y - 1, ny, g.per[1]), z)], o, A.Ox2[lin(g, x, bmap(y + 1, ny, g.per[1]), z)]);
    }
    if (A.f.use_Oy) {
        const R o = A.Oy2[i];
        A.Oy[i] = o + lap24<R>(A.Oy2[lin(g, x, bmap(y - 1, ny, g.per[1]), z)], o, A.Oy2[lin(g, x, bmap(y + 1, ny, g.per[1]), z)]) +
                  lap24<R>(A.Oy2[lin(g, x, y, bmap(z - 1, nz, g.per[2]))], o, A.Oy2[lin(g, x, y, bmap(z + 1, nz, g.per[2]))]);
    }
    if (A.f.use_Oz) {
        const R o = A.Oz2[i];
        A.Oz[i] = o + lap24<R>(A.Oz2[lin(g, bmap(x - 1, nx, g.per[0]), y, z)], o, A.Oz2[lin(g, bmap(x + 1, nx, g.per[0]), y, z)]) +
                  lap24<R>(A.Oz2[lin(g, x, y, bmap(z - 1, nz, g.per[2]))], o, A.Oz2[lin(g, x, y, bmap(z + 1, nz, g.per[2]))]);
    }
}

// k_emf2 and k_emf in one pass for three-dimensional grids (every CT flag
// set): the corner EMFs O2 of src/fluxct.cpp:809-821 and their smoothing,
// 828-840, without the O2 round trip through memory.  A block of 512
// threads holds a 64 x 8 tile of columns (EMF_OX x EMF_OY owned columns and a
// one-column halo) and marches through EMF_CZ planes.  Per plane z:
//  * the in-plane CT fluxes of plane z + 1 (gsx, hsx with a 1|2-column halo
//    for their x stencils, fsy, hsy with a 1|2-row halo) were loaded one
//    plane ahead and go to LDS; gsz and fsz come from per-column registers
//    holding planes z .. z + 3;
//  * each tile position computes O2 at z + 1 into the other half of a
//    double-buffered LDS tile;
//  * each owned column smooths O2 at z: in-plane neighbours from LDS, z
//    neighbours (O2y, O2z at z - 1, z + 1) from its registers.
// Tile positions are unmapped coordinates; a position's CT values are those
// of its bmap image.  Periodic images are translations, so a halo position's
// O2 is the O2 of its image; at an open edge the image of the outside
// neighbour is the cell itself (bmap clamps), so owned cells there read their
// own O2 instead of the halo's.  The first two planes of a chunk (O2 at
// z0 - 1 and z0) come straight from memory through bmap, as in k_emf2.  The
// arithmetic is k_emf2's and k_emf's, term for term.
#ifndef HOWMHD_EMF_WX
#define HOWMHD_EMF_WX 64
#endif
#ifndef HOWMHD_EMF_WY
#define HOWMHD_EMF_WY 8
#endif
#ifndef HOWMHD_EMF_CZ
#define HOWMHD_EMF_CZ 32
#endif
constexpr int EMF_WX = HOWMHD_EMF_WX, EMF_WY = HOWMHD_EMF_WY, EMF_CZ = HOWMHD_EMF_CZ;
constexpr int EMF_OX = EMF_WX - 2, EMF_OY = EMF_WY - 2;
constexpr int EMF_THREADS = EMF_WX * EMF_WY;
static_assert(6 * (EMF_WX + EMF_WY) <= EMF_THREADS, "one stencil-halo cell per thread at most");

template <typename R>
struct EmfMarchArgs {
    const R *fsy, *fsz, *gsx, *gsz, *hsx, *hsy;
    R *Ox, *Oy, *Oz;
    Geom g;
};

template <typename R, int EX>
__global__ __launch_bounds__(EMF_THREADS) void k_emf_march(EmfMarchArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    constexpr int WX = EMF_WX, WY = EMF_WY;
    constexpr int NXH = 3 * WY, NYH = 3 * WX;  // stencil halo cells of the x and y arrays
    __shared__ R sgx[WY][WX + 3], shx[WY][WX + 3];  // column c holds position c - 1 (x stencils)
    __shared__ R sfy[WY + 3][WX], shy[WY + 3][WX];  // row r holds position r - 1 (y stencils)
    __shared__ R sO[2][3][WY][WX];                  // O2 of two planes
    const Geom& g = A.g;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int px = g.per[0], py = g.per[1], pz = g.per[2];
    const int x0 = blockIdx.x * EMF_OX, y0 = blockIdx.y * EMF_OY;
    const int zs = blockIdx.z * EMF_CZ, ze = min(zs + EMF_CZ, nz);
    const int tid = threadIdx.x;
    const int ex = tid % WX, ey = tid / WX;
    // image of tile position p (x: column, y: row); positions beyond the
    // stencil reach of any needed O2 are clamped into bmap's range
    auto imx = [&](int pos) { const int a = x0 - 1 + pos; return bmap(a > nx + 2 ? nx + 2 : a, nx, px); };
    auto imy = [&](int pos) { const int a = y0 - 1 + pos; return bmap(a > ny + 2 ? ny + 2 : a, ny, py); };
    const int xm = imx(ex), ym = imy(ey);
    const int ax = x0 - 1 + ex, ay = y0 - 1 + ey;
    const bool own = ex >= 1 && ex <= EMF_OX && ey >= 1 && ey <= EMF_OY && ax < nx && ay < ny;
    // in-plane neighbour positions of an owned cell (open edges: itself)
    const int exl = (ax == 0 && !px) ? ex : ex - 1, exr = (ax == nx - 1 && !px) ? ex : ex + 1;
    const int eyl = (ay == 0 && !py) ? ey : ey - 1, eyr = (ay == ny - 1 && !py) ? ey : ey + 1;
    // this thread's stencil-halo cell: (array, LDS row, LDS column, image)
    int hk = -1, hr = 0, hc = 0, hxm = 0, hym = 0;
    if (tid < 2 * NXH) {             // gsx / hsx: columns at positions -1, WX, WX + 1
        hk = tid < NXH ? 0 : 1;
        const int t = tid % NXH, col = t % 3, row = t / 3;
        const int pos = col == 0 ? -1 : WX - 1 + col;
        hr = row; hc = pos + 1; hxm = imx(pos); hym = imy(row);
    } else if (tid < 2 * NXH + 2 * NYH) {  // fsy / hsy: rows at positions -1, WY, WY + 1
        const int t0 = tid - 2 * NXH;
        hk = t0 < NYH ? 2 : 3;
        const int t = t0 % NYH, row = t / WX, col = t % WX;
        const int pos = row == 0 ? -1 : WY - 1 + row;
        hr = pos + 1; hc = col; hxm = imx(col); hym = imy(pos);
    }
    const R* harr = hk == 0 ? A.gsx : hk == 1 ? A.hsx : hk == 2 ? A.fsy : A.hsy;

    // O2 (k_emf2) of stored cell (x, y, z), straight from memory
    auto o2 = [&](int x, int y, int z, R& ox2, R& oy2, R& oz2) {
        const int xm1 = bmap(x - 1, nx, px), xp1 = bmap(x + 1, nx, px), xp2 = bmap(x + 2, nx, px);
        const int ym1 = bmap(y - 1, ny, py), yp1 = bmap(y + 1, ny, py), yp2 = bmap(y + 2, ny, py);
        const int zm1 = bmap(z - 1, nz, pz), zp1 = bmap(z + 1, nz, pz), zp2 = bmap(z + 2, nz, pz);
        const long long i = lin(g, x, y, z);
        ox2 = i4<R>(A.gsx[lin(g, xm1, y, z)], A.gsx[i], A.gsx[lin(g, xp1, y, z)], A.gsx[lin(g, xp2, y, z)]) -
              i4<R>(A.fsy[lin(g, x, ym1, z)], A.fsy[i], A.fsy[lin(g, x, yp1, z)], A.fsy[lin(g, x, yp2, z)]);
        oy2 = i4<R>(A.hsy[lin(g, x, ym1, z)], A.hsy[i], A.hsy[lin(g, x, yp1, z)], A.hsy[lin(g, x, yp2, z)]) -
              i4<R>(A.gsz[lin(g, x, y, zm1)], A.gsz[i], A.gsz[lin(g, x, y, zp1)], A.gsz[lin(g, x, y, zp2)]);
        oz2 = i4<R>(A.fsz[lin(g, x, y, zm1)], A.fsz[i], A.fsz[lin(g, x, y, zp1)], A.fsz[lin(g, x, y, zp2)]) -
              i4<R>(A.hsx[lin(g, xm1, y, z)], A.hsx[i], A.hsx[lin(g, xp1, y, z)], A.hsx[lin(g, xp2, y, z)]);
    };
    // loads of plane m (an image plane) for the in-plane stencils
    R cgx, chx, cfy, chy, chalo = R(0);
    auto load_plane = [&](int m) {
        const long long i = lin(g, xm, ym, m);
        cgx = A.gsx[i]; chx = A.hsx[i]; cfy = A.fsy[i]; chy = A.hsy[i];
        if (hk >= 0) chalo = harr[lin(g, hxm, hym, m)];
    };

    // prologue: O2 at z0 - 1 (y, z components) and z0, from memory
    R ymv, zmv, yc, zc;
    {
        R ox, unused;
        o2(xm, ym, bmap(zs - 1, nz, pz), unused, ymv, zmv);
        o2(xm, ym, zs, ox, yc, zc);
        sO[zs & 1][0][ey][ex] = ox;
        sO[zs & 1][1][ey][ex] = yc;
        sO[zs & 1][2][ey][ex] = zc;
    }
    // gsz, fsz of this column at planes z .. z + 3 for O2 at z + 1 (the
    // first entry is shifted out before use)
    const long long col = lin(g, xm, ym, 0);
    const long long plane = (long long)nx * ny;
    R gz[4], fz[4];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        const int m = bmap(zs + k - 1, nz, pz);
        gz[k] = A.gsz[col + m * plane];
        fz[k] = A.fsz[col + m * plane];
    }
    gz[0] = fz[0] = R(0);
    R gzn = A.gsz[col + bmap(zs + 3, nz, pz) * plane], fzn = A.fsz[col + bmap(zs + 3, nz, pz) * plane];
    load_plane(bmap(zs + 1, nz, pz));

    for (int z = zs; z < ze; ++z) {
        // in-plane fluxes of plane z + 1 to LDS; column window z .. z + 3
        sgx[ey][ex + 1] = cgx;
        shx[ey][ex + 1] = chx;
        sfy[ey + 1][ex] = cfy;
        shy[ey + 1][ex] = chy;
        if (hk == 0) sgx[hr][hc] = chalo;
        else if (hk == 1) shx[hr][hc] = chalo;
        else if (hk == 2) sfy[hr][hc] = chalo;
        else if (hk == 3) shy[hr][hc] = chalo;
#pragma unroll
        for (int k = 0; k < 3; ++k) { gz[k] = gz[k + 1]; fz[k] = fz[k + 1]; }
        gz[3] = gzn; fz[3] = fzn;
        // loads of the next plane, in flight across the barriers below
        if (z + 1 < ze) {
            load_plane(bmap(z + 2, nz, pz));
            const int m = bmap(z + 4, nz, pz);
            gzn = A.gsz[col + m * plane];
            fzn = A.fsz[col + m * plane];
        }
        __syncthreads();
        // O2 at z + 1 at this tile position (open top edge: the image of
        // plane nz is plane nz - 1, whose O2 is this iteration's)
        R yn, zn;
        const int cb = (z + 1) & 1;
        if (z + 1 == nz && !pz) {
            yn = yc; zn = zc;
        } else {
            const R ox2 = i4<R>(sgx[ey][ex], sgx[ey][ex + 1], sgx[ey][ex + 2], sgx[ey][ex + 3]) -
                          i4<R>(sfy[ey][ex], sfy[ey + 1][ex], sfy[ey + 2][ex], sfy[ey + 3][ex]);
            yn = i4<R>(shy[ey][ex], shy[ey + 1][ex], shy[ey + 2][ex], shy[ey + 3][ex]) -
                 i4<R>(gz[0], gz[1], gz[2], gz[3]);
            zn = i4<R>(fz[0], fz[1], fz[2], fz[3]) -
                 i4<R>(shx[ey][ex], shx[ey][ex + 1], shx[ey][ex + 2], shx[ey][ex + 3]);
            sO[cb][0][ey][ex] = ox2;
            sO[cb][1][ey][ex] = yn;
            sO[cb][2][ey][ex] = zn;
        }
        // smoothing of O2 at z (k_emf)
        if (own) {
            const int c = z & 1;
            const long long i = lin(g, xm, ym, z);
            const R ox = sO[c][0][ey][ex], oy = yc, oz = zc;
            st_nt(&A.Ox[i], ox + lap24<R>(sO[c][0][ey][exl], ox, sO[c][0][ey][exr]) +
                                lap24<R>(sO[c][0][eyl][ex], ox, sO[c][0][eyr][ex]));
            st_nt(&A.Oy[i], oy + lap24<R>(sO[c][1][eyl][ex], oy, sO[c][1][eyr][ex]) + lap24<R>(ymv, oy, yn));
            st_nt(&A.Oz[i], oz + lap24<R>(sO[c][2][ey][exl], oz, sO[c][2][ey][exr]) + lap24<R>(zmv, oz, zn));
        }
        __syncthreads();
        ymv = yc; zmv = zc;
        yc = yn; zc = zn;
    }
}

struct StageCoef {
    double k1, k3, c4, c5;  // c4 = k4/k1, c5 = k5/k1 (stage 5 only)
    double dtdx, dtdy, dtdz;
    int stage5;
};

template <typename R>
struct FaceArgs {
    const R *Ox, *Oy, *Oz;
    const double* b0;   // stage-0 faces (3 arrays of ncell)
    const double* b2;   // stage 2 (stage 5 only)
    const double* b3;   // stage 3 (stage 5 only)
    const double* bp;   // previous stage
    double* bo;         // this stage
    Geom g;
    CtFlags f;
    StageCoef s;
};

// Staggered 4th-order difference aa*(o0-om1) + bb*(op1-om2) + cc*(op2-om3),
// aa = 9/8, bb = -1/24, cc = 0 (src/fluxct.cpp:848, 861-874), in FP64.
__device__ __forceinline__ double d4(double om3, double om2, double om1, double o0, double op1, double op2) {
    const double aa = 9.0 / 8.0, bb = -1.0 / 24.0, cc = 0.0;
    return aa * (o0 - om1) + bb * (op1 - om2) + cc * (op2 - om3);
}

// fluxct_update_bfield, src/fluxct.cpp:850-877, with the stage initialisation
// of copy_bfield_stage / prepare_bfield_stage5 (111-135) folded in.
// Block = FACE_TX x FACE_TY cells of one z plane.  The in-plane stencils
// (Ox along x and y, Oy along y, Oz along x) are read from LDS copies of the
// tile with its 3-cell halo (each value read once per block instead of six
// times through L2); the z stencils of Oy and Oz are read from memory.
constexpr int FACE_TX = 32, FACE_TY = 8;
template <typename R, int EX>
__global__ __launch_bounds__(256) void k_faces(FaceArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    constexpr int HX = FACE_TX + 5, HY = FACE_TY + 5;
    __shared__ R sOx[HY][HX];       // rows y0-3 .. y0+TY+1, columns x0-3 .. x0+TX+1
    __shared__ R sOy[HY][FACE_TX];  // rows y0-3 .. y0+TY+1
    __shared__ R sOz[FACE_TY][HX];  // columns x0-3 .. x0+TX+1
    const Geom& g = A.g;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int px = g.per[0], py = g.per[1], pz = g.per[2];
    const int x0 = blockIdx.x * FACE_TX, y0 = blockIdx.y * FACE_TY, z = blockIdx.z;
    const int tid = threadIdx.x;
    const CtFlags& f = A.f;
    // halo coordinates beyond the stencil reach of any valid cell are clamped
    // into the range bmap covers (those entries are never read)
    auto hx = [&](int c) { const int a = x0 - 3 + c; return bmap(a > nx + 2 ? nx + 2 : a, nx, px); };
    auto hy = [&](int r) { const int a = y0 - 3 + r; return bmap(a > ny + 2 ? ny + 2 : a, ny, py); };
    if (f.active) {
        if (f.use_Ox)
            for (int k = tid; k < HY * HX; k += 256) sOx[k / HX][k % HX] = A.Ox[lin(g, hx(k % HX), hy(k / HX), z)];
        if (f.use_Oy)
            for (int k = tid; k < HY * FACE_TX; k += 256) {
                const int c = k % FACE_TX, xx = x0 + c < nx ? x0 + c : nx - 1;
                sOy[k / FACE_TX][c] = A.Oy[lin(g, xx, hy(k / FACE_TX), z)];
            }
        if (f.use_Oz)
            for (int k = tid; k < FACE_TY * HX; k += 256) {
                const int r = k / HX, yy = y0 + r < ny ? y0 + r : ny - 1;
                sOz[r][k % HX] = A.Oz[lin(g, hx(k % HX), yy, z)];
            }
    }
    __syncthreads();
    const int tx = tid % FACE_TX, ty = tid / FACE_TX;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= nx || y >= ny) return;
    const long long i = lin(g, x, y, z);
    const StageCoef& s = A.s;
    double binit[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const long long o = (long long)c * g.ncell + i;
        binit[c] = s.stage5 ? A.b0[o] + s.c4 * A.b2[o] + s.c5 * A.b3[o] : A.b0[o];
    }
    auto dz_of = [&](const R* O) -> double {
        auto at = [&](int zz) -> double { return (double)O[lin(g, x, y, zz)]; };
        return d4(at(bmap(z - 3, nz, pz)), at(bmap(z - 2, nz, pz)), at(bmap(z - 1, nz, pz)), at(z),
                  at(bmap(z + 1, nz, pz)), at(bmap(z + 2, nz, pz)));
    };
    auto col6 = [&](const R (*t)[HX], int c) -> double {  // rows y-3..y+2 of column c
        return d4((double)t[ty][c], (double)t[ty + 1][c], (double)t[ty + 2][c], (double)t[ty + 3][c],
                  (double)t[ty + 4][c], (double)t[ty + 5][c]);
    };
    auto row6 = [&](const R* r) -> double {  // columns x-3..x+2 of a tile row
        return d4((double)r[tx], (double)r[tx + 1], (double)r[tx + 2], (double)r[tx + 3], (double)r[tx + 4],
                  (double)r[tx + 5]);
    };
    double bx = binit[0], by = binit[1], bz = binit[2];
    if (f.active) {
        if (f.use_Ox || f.use_Oz) {
            double dOx = 0.0, dOz = 0.0;
            if (f.use_Ox) dOx = col6(sOx, tx + 3);
            if (f.use_Oz) dOz = dz_of(A.Oz);
            bx = s.k1 * binit[0] + s.k3 * A.bp[i] - s.dtdy * dOx + s.dtdz * dOz;
        }
        if (f.use_Oy || f.use_Ox) {
            double dOy = 0.0, dOx = 0.0;
            if (f.use_Oy) dOy = dz_of(A.Oy);
            if (f.use_Ox) dOx = row6(sOx[ty + 3]);
            by = s.k1 * binit[1] + s.k3 * A.bp[g.ncell + i] - s.dtdz * dOy + s.dtdx * dOx;
        }
        if (f.use_Oz || f.use_Oy) {
            double dOz = 0.0, dOy = 0.0;
            if (f.use_Oz) dOz = row6(sOz[ty]);
            if (f.use_Oy)
                dOy = d4((double)sOy[ty][tx], (double)sOy[ty + 1][tx], (double)sOy[ty + 2][tx], (double)sOy[ty + 3][tx],
                         (double)sOy[ty + 4][tx], (double)sOy[ty + 5][tx]);
            bz = s.k1 * binit[2] + s.k3 * A.bp[2 * g.ncell + i] - s.dtdx * dOz + s.dtdy * dOy;
        }
    }
    st_nt(&A.bo[i], bx);
    st_nt(&A.bo[g.ncell + i], by);
    st_nt(&A.bo[2 * g.ncell + i], bz);
}

// k_faces for three-dimensional grids (every CT flag set), marching a
// FACE_TX x FACE_TY column tile through FACE_CZ planes: the in-plane EMF
// tiles of each plane (the LDS tiles of k_faces) and the face values of the
// next plane are loaded one plane ahead, and each column keeps Oy and Oz of
// planes z - 3 .. z + 2 (images through bmap) in registers for the z
// differences.  The arithmetic is k_faces', term for term.
#ifndef HOWMHD_FACE_CZ
#define HOWMHD_FACE_CZ 32
#endif
constexpr int FACE_CZ = HOWMHD_FACE_CZ;
// 3 wavefronts per SIMD (136 VGPRs, no scratch in FP64; 4 spills 4 values and
// runs 4% slower)
#ifndef HOWMHD_FACE_WAVES
#define HOWMHD_FACE_WAVES 3
#endif
template <typename R, int EX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HOWMHD_FACE_WAVES))) void k_faces_march(FaceArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    constexpr int HX = FACE_TX + 5, HY = FACE_TY + 5;
    constexpr int NOX = HY * HX, NOY = HY * FACE_TX, NOZ = FACE_TY * HX, NF = NOX + NOY + NOZ;
    constexpr int NFT = (NF + 255) / 256;  // tile values per thread
    // sOx[HY][HX]: rows y0-3 .. y0+TY+1, columns x0-3 .. x0+TX+1; then
    // sOy[HY][TX] (rows y0-3 ..), sOz[TY][HX] (columns x0-3 ..)
    __shared__ R sbuf[NF];
    const R(*sOx)[HX] = reinterpret_cast<const R(*)[HX]>(sbuf);
    const R(*sOy)[FACE_TX] = reinterpret_cast<const R(*)[FACE_TX]>(sbuf + NOX);
    const R(*sOz)[HX] = reinterpret_cast<const R(*)[HX]>(sbuf + NOX + NOY);
    const Geom& g = A.g;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int px = g.per[0], py = g.per[1], pz = g.per[2];
    const int x0 = blockIdx.x * FACE_TX, y0 = blockIdx.y * FACE_TY;
    const int zs = blockIdx.z * FACE_CZ, ze = min(zs + FACE_CZ, nz);
    const int tid = threadIdx.x;
    const long long plane = (long long)nx * ny;
    auto hx = [&](int c) { const int a = x0 - 3 + c; return bmap(a > nx + 2 ? nx + 2 : a, nx, px); };
    auto hy = [&](int r) { const int a = y0 - 3 + r; return bmap(a > ny + 2 ? ny + 2 : a, ny, py); };
    // this thread's tile values: source column (array and in-plane offset)
    const R* src[NFT];
#pragma unroll
    for (int k = 0; k < NFT; ++k) {
        const int e = tid + 256 * k;
        src[k] = nullptr;
        if (e < NOX) {
            src[k] = A.Ox + (long long)hy(e / HX) * nx + hx(e % HX);
        } else if (e < NOX + NOY) {
            const int t = e - NOX, c = t % FACE_TX, xx = x0 + c < nx ? x0 + c : nx - 1;
            src[k] = A.Oy + (long long)hy(t / FACE_TX) * nx + xx;
        } else if (e < NF) {
            const int t = e - NOX - NOY, r = t / HX, yy = y0 + r < ny ? y0 + r : ny - 1;
            src[k] = A.Oz + (long long)yy * nx + hx(t % HX);
        }
    }
    const int tx = tid % FACE_TX, ty = tid / FACE_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < nx && y < ny;
    const long long col = (long long)(valid ? y : 0) * nx + (valid ? x : 0);
    const StageCoef& s = A.s;
    R tv[NFT];
    auto load_tile = [&](int z) {
#pragma unroll
        for (int k = 0; k < NFT; ++k)
            if (src[k]) tv[k] = src[k][(long long)z * plane];
    };
    // face values of a plane: b0 (+ b2, b3 at stage 5) and the previous stage
    double fb0[3], fb2[3], fb3[3], fbp[3];
    auto load_faces = [&](int z) {
        const long long i = col + (long long)z * plane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long o = (long long)c * g.ncell + i;
            fb0[c] = ld_nt(&A.b0[o]);
            fbp[c] = ld_nt(&A.bp[o]);
            if (s.stage5) { fb2[c] = ld_nt(&A.b2[o]); fb3[c] = ld_nt(&A.b3[o]); }
            else { fb2[c] = 0.0; fb3[c] = 0.0; }
        }
    };
    R wy[6], wz[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const long long o = col + (long long)bmap(zs - 3 + k, nz, pz) * plane;
        wy[k] = A.Oy[o];
        wz[k] = A.Oz[o];
    }
    load_tile(zs);
    load_faces(zs);
    for (int z = zs; z < ze; ++z) {
#pragma unroll
        for (int k = 0; k < NFT; ++k)
            if (src[k]) sbuf[tid + 256 * k] = tv[k];
        double binit[3], bp[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            binit[c] = s.stage5 ? fb0[c] + s.c4 * fb2[c] + s.c5 * fb3[c] : fb0[c];
            bp[c] = fbp[c];
        }
        R ny_ = R(0), nz_ = R(0);
        if (z + 1 < ze) {  // next plane, in flight across the barriers below
            load_tile(z + 1);
            load_faces(z + 1);
            const long long o = col + (long long)bmap(z + 3, nz, pz) * plane;
            ny_ = A.Oy[o];
            nz_ = A.Oz[o];
        }
        __syncthreads();
        if (valid) {
            const long long i = col + (long long)z * plane;
            auto col6 = [&](const R (*t)[HX], int c) -> double {
                return d4((double)t[ty][c], (double)t[ty + 1][c], (double)t[ty + 2][c], (double)t[ty + 3][c],
                          (double)t[ty + 4][c], (double)t[ty + 5][c]);
            };
            auto row6 = [&](const R* r) -> double {
                return d4((double)r[tx], (double)r[tx + 1], (double)r[tx + 2], (double)r[tx + 3], (double)r[tx + 4],
                          (double)r[tx + 5]);
            };
            const double dzOy = d4((double)wy[0], (double)wy[1], (double)wy[2], (double)wy[3], (double)wy[4],
                                   (double)wy[5]);
            const double dzOz = d4((double)wz[0], (double)wz[1], (double)wz[2], (double)wz[3], (double)wz[4],
                                   (double)wz[5]);
            const double bx = s.k1 * binit[0] + s.k3 * bp[0] - s.dtdy * col6(sOx, tx + 3) + s.dtdz * dzOz;
            const double by = s.k1 * binit[1] + s.k3 * bp[1] - s.dtdz * dzOy + s.dtdx * row6(sOx[ty + 3]);
            const double dyOy = d4((double)sOy[ty][tx], (double)sOy[ty + 1][tx], (double)sOy[ty + 2][tx],
                                   (double)sOy[ty + 3][tx], (double)sOy[ty + 4][tx], (double)sOy[ty + 5][tx]);
            const double bz = s.k1 * binit[2] + s.k3 * bp[2] - s.dtdx * row6(sOz[ty]) + s.dtdy * dyOy;
            st_nt(&A.bo[i], bx);
            st_nt(&A.bo[g.ncell + i], by);
            st_nt(&A.bo[2 * g.ncell + i], bz);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 5; ++k) { wy[k] = wy[k + 1]; wz[k] = wz[k + 1]; }
        wy[5] = ny_;
        wz[5] = nz_;
    }
}

// CFL speeds of one cell (src/tstep.cpp:28-83): the running maxima of
// |v_d| + c_fast,d from the conserved values, with the reference's rounding
// in every unit (no contraction here).
__device__ __forceinline__ void cfl_cell(const double* q, const Phys& ph, double& m0, double& m1, double& m2) {
#pragma clang fp contract(off)
    const double gam = ph.gam;
    const double DD = q[0], Mx = q[1], My = q[2], Mz = q[3], Bx = q[4], By = q[5], Bz = q[6], EE = q[7];
    const double rho = fmax(DD, ph.rhomin);
    const double vx = Mx / rho, vy = My / rho, vz = Mz / rho;
    const double vv2 = vx * vx + vy * vy + vz * vz;
    const double BB2 = Bx * Bx + By * By + Bz * Bz;
    const double pg = fmax(ph.pgmin, (gam - 1.0) * (EE - 0.5 * (rho * vv2 + BB2)));
    const double bbn2 = BB2 / rho;
    const double cs2 = fmax(0.0, gam * pg / rho);
    const double bnx2 = Bx * Bx / rho, bny2 = By * By / rho, bnz2 = Bz * Bz / rho;
    const double rootx = sqrt(fmax(0.0, (bbn2 + cs2) * (bbn2 + cs2) - 4.0 * bnx2 * cs2));
    const double rooty = sqrt(fmax(0.0, (bbn2 + cs2) * (bbn2 + cs2) - 4.0 * bny2 * cs2));
    const double rootz = sqrt(fmax(0.0, (bbn2 + cs2) * (bbn2 + cs2) - 4.0 * bnz2 * cs2));
    const double lfx = sqrt(fmax(0.0, 0.5 * (bbn2 + cs2 + rootx)));
    const double lfy = sqrt(fmax(0.0, 0.5 * (bbn2 + cs2 + rooty)));
    const double lfz = sqrt(fmax(0.0, 0.5 * (bbn2 + cs2 + rootz)));
    m0 = fmax(m0, fabs(vx) + lfx);
    m1 = fmax(m1, fabs(vy) + lfy);
    m2 = fmax(m2, fabs(vz) + lfz);
}

// Maxima of m0..m2 over the 256 threads of the block (all must call it), in
// out[0..2] of thread 0.
__device__ __forceinline__ void block_max3(double m0, double m1, double m2, double* out) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, off, 64));
        m1 = fmax(m1, __shfl_xor(m1, off, 64));
        m2 = fmax(m2, __shfl_xor(m2, off, 64));
    }
    __shared__ double red[3][4];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { red[0][w] = m0; red[1][w] = m1; red[2][w] = m2; }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[d] = fmax(fmax(red[d][0], red[d][1]), fmax(red[d][2], red[d][3]));
}

// Counts of the lanes with a (b) set over the 256 threads of the block (all
// must call it), in out[0..1].
__device__ __forceinline__ void block_count2(bool a, bool b, unsigned* out) {
    __shared__ unsigned cnt[2][4];
    const int w = threadIdx.x / 64;
    const unsigned ca = (unsigned)__popcll(__ballot(a)), cb = (unsigned)__popcll(__ballot(b));
    if ((threadIdx.x & 63) == 0) { cnt[0][w] = ca; cnt[1][w] = cb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = cnt[0][0] + cnt[0][1] + cnt[0][2] + cnt[0][3];
        out[1] = cnt[1][0] + cnt[1][1] + cnt[1][2] + cnt[1][3];
    }
}

template <typename R>
struct CombineArgs {
    const double* q0;
    const double* q2;   // stage 5 only
    const double* q3;   // stage 5 only
    const double* qp;   // previous stage
    double* qo;         // this stage
    const double* bo;   // faces of this stage
    const R *dFx, *dFy, *dFz;
    Geom g;
    CtFlags f;
    StageCoef s;
    Phys ph;
    // stage 5: per-block CFL maxima of the new q(0) (3 doubles per block),
    // which the next tstep reduces instead of reading q(0) again, and the
    // per-block counts of cells whose density and pressure floors prot
    // applied (2 per block, read by the diagnostics); else null
    double* cfl;
    unsigned* floors;
};

// 6th-order face -> centre interpolation of src/fluxct.cpp:891-896:
// (3 b(-3) - 25 b(-2) + 150 b(-1) + 150 b(0) - 25 b(1) + 3 b(2)) / 256
__device__ __forceinline__ double i6(double a, double b, double c, double d, double e, double f) {
    return div_c<double>(3.0 * a - 25.0 * b + 150.0 * c + 150.0 * d - 25.0 * e + 3.0 * f, 256.0, 1.0 / 256.0);
}

// Stage combination of one cell (x, y, z; linear index i): x_sweep's store
// q = k1*q + k3*q_prev + dX (src/ssprk.cpp:415-430), then += dY, += dZ
// (1075-1094, 1166-1185); face -> centre B (fluxct.cpp:882-897); prot
// (prot.cpp:10-51).  v receives the new q of the cell; rho_floor / p_floor
// whether prot raised rho / p (a value below the floor or not a number:
// fmax returns the floor for both).
template <typename R>
__device__ __forceinline__ void combine_cell(const CombineArgs<R>& A, long long i, int x, int y, int z, double* v,
                                             bool& rho_floor, bool& p_floor) {
    const Geom& g = A.g;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const StageCoef& s = A.s;
    // With CT active the cell-centred B (m = 4..6) is replaced by the face
    // interpolation below, so its flux update is not formed (its dF
    // components are not stored either).
    const bool ct = A.f.active;
    auto upd = [&](int m) { return !(ct && m >= 4 && m <= 6); };
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const long long o = (long long)m * g.ncell + i;
        v[m] = 0.0;
        if (upd(m)) v[m] = s.stage5 ? ld_nt(&A.q0[o]) + s.c4 * ld_nt(&A.q2[o]) + s.c5 * ld_nt(&A.q3[o]) : ld_nt(&A.q0[o]);
    }
    // slot of each q variable in the dF of each sweep (-1: the normal field)
    // x: (rho, Mx, My, Mz, Bx, By, Bz, E) -> (0,1,2,3,-,4,5,6)
    // y: rho->0, Mx->3, My->1, Mz->2, Bx->5, By->-, Bz->4, E->6
    // z: rho->0, Mx->2, My->3, Mz->1, Bx->4, By->5, Bz->-, E->6
    constexpr int sx[8] = {0, 1, 2, 3, -1, 4, 5, 6};
    constexpr int sy[8] = {0, 3, 1, 2, 5, -1, 4, 6};
    constexpr int sz[8] = {0, 2, 3, 1, 4, 5, -1, 6};
    if (nx > 1) {
        const int tl = (x == 0 && g.per[0]) ? nx : x, tr = x + 1;
        const double dt = s.dtdx;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (!upd(m)) continue;
            double d = 0.0;
            if (sx[m] >= 0)
                d = -dt * ((double)A.dFx[dF_index<0>(g, sx[m], tr, x, y, z)] - (double)A.dFx[dF_index<0>(g, sx[m], tl, x, y, z)]);
            v[m] = s.k1 * v[m] + s.k3 * ld_nt(&A.qp[(long long)m * g.ncell + i]) + d;
        }
    }
    if (ny > 1) {
        const int tl = (y == 0 && g.per[1]) ? ny : y, tr = y + 1;
        const double dt = s.dtdy;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (!upd(m)) continue;
            double d = 0.0;
            if (sy[m] >= 0)
                d = -dt * ((double)A.dFy[dF_index<1>(g, sy[m], tr, x, y, z)] - (double)A.dFy[dF_index<1>(g, sy[m], tl, x, y, z)]);
            v[m] = v[m] + d;
        }
    }
    if (nz > 1) {
        const int tl = (z == 0 && g.per[2]) ? nz : z, tr = z + 1;
        const double dt = s.dtdz;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (!upd(m)) continue;
            double d = 0.0;
            if (sz[m] >= 0)
                d = -dt * ((double)A.dFz[dF_index<2>(g, sz[m], tr, x, y, z)] - (double)A.dFz[dF_index<2>(g, sz[m], tl, x, y, z)]);
            v[m] = v[m] + d;
        }
    }
    if (A.f.active) {
        const double* bx = A.bo;
        const double* by = A.bo + g.ncell;
        const double* bz = A.bo + 2 * g.ncell;
        if (A.f.use_x) {
            const int p = g.per[0];
            v[4] = i6(bx[lin(g, bmap(x - 3, nx, p), y, z)], bx[lin(g, bmap(x - 2, nx, p), y, z)], bx[lin(g, bmap(x - 1, nx, p), y, z)],
                      bx[i], bx[lin(g, bmap(x + 1, nx, p), y, z)], bx[lin(g, bmap(x + 2, nx, p), y, z)]);
        } else {
            v[4] = bx[i];
        }
        if (A.f.use_y) {
            const int p = g.per[1];
            v[5] = i6(by[lin(g, x, bmap(y - 3, ny, p), z)], by[lin(g, x, bmap(y - 2, ny, p), z)], by[lin(g, x, bmap(y - 1, ny, p), z)],
                      by[i], by[lin(g, x, bmap(y + 1, ny, p), z)], by[lin(g, x, bmap(y + 2, ny, p), z)]);
        } else {
            v[5] = by[i];
        }
        if (A.f.use_z) {
            const int p = g.per[2];
            v[6] = i6(bz[lin(g, x, y, bmap(z - 3, nz, p))], bz[lin(g, x, y, bmap(z - 2, nz, p))], bz[lin(g, x, y, bmap(z - 1, nz, p))],
                      bz[i], bz[lin(g, x, y, bmap(z + 1, nz, p))], bz[lin(g, x, y, bmap(z + 2, nz, p))]);
        } else {
            v[6] = bz[i];
        }
    }
    // prot
    const double gam = A.ph.gam;
    rho_floor = !(v[0] >= A.ph.rhomin);
    double rho = fmax(v[0], A.ph.rhomin);
    const double vx = v[1] / rho, vy = v[2] / rho, vz = v[3] / rho;
    const double vv2 = vx * vx + vy * vy + vz * vz;
    const double BB2 = v[4] * v[4] + v[5] * v[5] + v[6] * v[6];
    double pg = (gam - 1.0) * (v[7] - 0.5 * (rho * vv2 + BB2));
    p_floor = !(pg >= A.ph.pgmin);
    pg = fmax(pg, A.ph.pgmin);
    v[7] = pg / (gam - 1.0) + 0.5 * (rho * vv2 + BB2);
    v[0] = rho;
}

// The stage combination over every cell (one thread per cell); at stage 5
// also the per-block CFL maxima and floor counts.
template <typename R, int EX>
__global__ __launch_bounds__(256) void k_combine(CombineArgs<R> A) {
    static_assert(EX == HOWMHD_EXACT, "kernel instantiated in a unit of the other rounding form");
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const Geom& g = A.g;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;  // CFL maxima (stage 5)
    bool rho_floor = false, p_floor = false;  // prot raised rho / p (stage 5)
    if (i < g.ncell) {
        int x, y, z;
        cell_of(g, i, x, y, z);
        double v[8];
        combine_cell<R>(A, i, x, y, z, v, rho_floor, p_floor);
#pragma unroll
        for (int m = 0; m < 8; ++m) st_nt(&A.qo[(long long)m * g.ncell + i], v[m]);
        if (A.cfl) cfl_cell(v, A.ph, m0, m1, m2);
    }
    if (A.cfl) {
        block_max3(m0, m1, m2, A.cfl + 3 * (long long)blockIdx.x);
        block_count2(rho_floor, p_floor, A.floors + 2 * (long long)blockIdx.x);
    }
}

}  // namespace howmhd
