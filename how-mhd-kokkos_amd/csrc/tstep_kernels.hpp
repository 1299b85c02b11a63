// Kernels of the per-step calls (tstep and the face initialisation) and of
// the host transfers, built once, in the plain FP64 unit
// (-ffp-contract=off): the CFL maximum and dt are the reference's bit for bit
// for every form.
#pragma once

#include "kernels.hpp"

namespace howmhd {

// ---------------------------------------------------------------- CFL (tstep)
// Per-direction maximum of |v_d| + c_fast,d over active cells of q(0)
// (src/tstep.cpp:28-83).  Non-negative doubles order like their bit patterns,
// so the block maxima are combined with a 64-bit atomicMax.
__global__ __launch_bounds__(256) void k_cfl(const double* q, Geom g, Phys ph, unsigned long long* out) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < g.ncell; i += (long long)gridDim.x * 256) {
        double c[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) c[m] = q[m * g.ncell + i];
        cfl_cell(c, ph, m0, m1, m2);
    }
    __shared__ double bm[3];
    block_max3(m0, m1, m2, bm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d) atomicMax(&out[d], (unsigned long long)__double_as_longlong(bm[d]));
}

// The same maxima from the per-block maxima that the stage-5 combination
// of the last step wrote (n blocks x 3).
__global__ __launch_bounds__(256) void k_cfl_fold(const double* part, long long n, unsigned long long* out) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (long long b = (long long)blockIdx.x * 256 + threadIdx.x; b < n; b += (long long)gridDim.x * 256) {
        m0 = fmax(m0, part[3 * b]);
        m1 = fmax(m1, part[3 * b + 1]);
        m2 = fmax(m2, part[3 * b + 2]);
    }
    __shared__ double bm[3];
    block_max3(m0, m1, m2, bm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d) atomicMax(&out[d], (unsigned long long)__double_as_longlong(bm[d]));
}

// The same maxima of a q(0) that the caller holds on the device in the
// reference layout: 8 arrays of (nx+6)(ny+6)(nz+6) values, three ghost layers
// per side, k fastest (ghosts are not read).  At most CFL_REF_BLOCKS blocks
// walk the active cells with z fastest, so that a wavefront reads consecutive
// k; block b leaves its maximum triple in part[3b..3b+2], and k_cfl_ref_fold
// folds the triples of the nb blocks into part[3 CFL_REF_BLOCKS..+2].  Every
// block writes its triple, so nothing has to be cleared before a launch.
constexpr int CFL_REF_BLOCKS = 1024;

__global__ __launch_bounds__(256) void k_cfl_ref(const double* q, Geom g, Phys ph, double* part) {
    const long long nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const long long n6y = ny + 6, n6z = nz + 6;
    const long long vstride = (nx + 6) * n6y * n6z;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < g.ncell; i += (long long)gridDim.x * 256) {
        const long long z = i % nz, t = i / nz;
        const long long y = t % ny, x = t / ny;
        const long long o = ((x + 3) * n6y + y + 3) * n6z + z + 3;
        double c[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) c[m] = q[m * vstride + o];
        cfl_cell(c, ph, m0, m1, m2);
    }
    block_max3(m0, m1, m2, part + 3 * (long long)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cfl_ref_fold(double* part, int nb) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        m0 = fmax(m0, part[3 * b]);
        m1 = fmax(m1, part[3 * b + 1]);
        m2 = fmax(m2, part[3 * b + 2]);
    }
    block_max3(m0, m1, m2, part + 3 * CFL_REF_BLOCKS);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Adds the per-wavefront counters of the routed sweeps (n slots of
// (flagged, total, wavefronts flagged, wavefronts total)) to the four totals
// and to (flagged, total) of the step in last[0..1] (cleared before the
// launch), and clears the slots.  Grid-stride loop over the slots; one atomic
// per total and block.
__global__ __launch_bounds__(256) void k_fold_wave_stats(uint4* w, long long n, unsigned long long* tot,
                                                         unsigned long long* last) {
    unsigned long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const uint4 v = w[i];
        s0 += v.x; s1 += v.y; s2 += v.z; s3 += v.w;
        w[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_down(s0, o, 64);
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
        s3 += __shfl_down(s3, o, 64);
    }
    __shared__ unsigned long long red[4][4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wv] = s0; red[1][wv] = s1; red[2][wv] = s2; red[3][wv] = s3; }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        const unsigned long long v = red[k][0] + red[k][1] + red[k][2] + red[k][3];
        atomicAdd(&tot[k], v);
        if (k < 2) atomicAdd(&last[k], v);
    }
}

// ------------------------------------------------------------- diagnostics
// Diagnostics of stage-0 q and faces (the divB field of src/output.cpp:154-188,
// 314-347 and the CT face divergence), reduced on the device in FP64.  A
// fixed grid of DIAG_BLOCKS blocks walks the cells; each block leaves its
// partials (DIAG_NV values) and k_diag_fold combines them in block order,
// so the sums do not depend on scheduling.
constexpr int DIAG_BLOCKS = 1024;
// sums of the 8 conserved variables, max |div_CT|, max |cell divB|,
// -min rho, -min p, and the floor counts of the last stage-5 prot
enum { DG_SUM = 0, DG_DIVCT = 8, DG_DIVC, DG_NRHO, DG_NP, DG_NFRHO, DG_NFP, DIAG_NV };

struct DiagArgs {
    const double* q;          // stage-0 q, 8 arrays of ncell
    const double* b;          // stage-0 faces, 3 arrays of ncell
    Geom g;
    double gam, dx, dy, dz;
    const unsigned* floors;   // 2 per block of the stage-5 combination, or null
    long long nfloor;         // blocks of that combination
    double* part;             // DIAG_BLOCKS x DIAG_NV
};

// Sum, or maximum, of v over the 256 threads of the block, in thread 0.
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double* red4) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = MAX ? fmax(v, o) : v + o;
    }
    const int w = threadIdx.x / 64;
    __syncthreads();  // red4 is free
    if ((threadIdx.x & 63) == 0) red4[w] = v;
    __syncthreads();
    if (MAX) return fmax(fmax(red4[0], red4[1]), fmax(red4[2], red4[3]));
    return (red4[0] + red4[1]) + (red4[2] + red4[3]);
}

__global__ __launch_bounds__(256) void k_diag(DiagArgs A) {
    const Geom& g = A.g;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int px = g.per[0], py = g.per[1], pz = g.per[2];
    const long long nc = g.ncell;
    double sum[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double mct = 0.0, mcell = 0.0, nrho = -INFINITY, np = -INFINITY, frho = 0.0, fp = 0.0;
    const double aa = 9.0 / 8.0, bb = -1.0 / 24.0;
    const double* bx = A.b;
    const double* by = A.b + nc;
    const double* bz = A.b + 2 * nc;
    const double* Bx = A.q + 4 * nc;
    const double* By = A.q + 5 * nc;
    const double* Bz = A.q + 6 * nc;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nc; i += (long long)gridDim.x * 256) {
        int x, y, z;
        cell_of(g, i, x, y, z);
        double v[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            v[m] = A.q[m * nc + i];
            sum[m] += v[m];
        }
        // pressure as src/output.cpp:300-312 forms it
        const double vx = v[1] / v[0], vy = v[2] / v[0], vz = v[3] / v[0];
        const double vv2 = vx * vx + vy * vy + vz * vz;
        const double BB2 = v[4] * v[4] + v[5] * v[5] + v[6] * v[6];
        const double pg = (A.gam - 1.0) * (v[7] - 0.5 * v[0] * vv2 - 0.5 * BB2);
        nrho = fmax(nrho, -v[0]);
        np = fmax(np, -pg);
        // CT face divergence: D4 of src/fluxct.cpp:848 along each axis
        double dv = 0.0;
        if (nx > 1)
            dv += (aa * (bx[i] - bx[lin(g, bmap(x - 1, nx, px), y, z)]) +
                   bb * (bx[lin(g, bmap(x + 1, nx, px), y, z)] - bx[lin(g, bmap(x - 2, nx, px), y, z)])) / A.dx;
        if (ny > 1)
            dv += (aa * (by[i] - by[lin(g, x, bmap(y - 1, ny, py), z)]) +
                   bb * (by[lin(g, x, bmap(y + 1, ny, py), z)] - by[lin(g, x, bmap(y - 2, ny, py), z)])) / A.dy;
        if (nz > 1)
            dv += (aa * (bz[i] - bz[lin(g, x, y, bmap(z - 1, nz, pz))]) +
                   bb * (bz[lin(g, x, y, bmap(z + 1, nz, pz))] - bz[lin(g, x, y, bmap(z - 2, nz, pz))])) / A.dz;
        mct = fmax(mct, fabs(dv));
        // 2nd-order cell-centred divB, src/output.cpp:314-347: central
        // difference, one-sided where the open edge clamps a neighbour
        double dc = 0.0;
        if (nx > 1) {
            const int xm = bmap(x - 1, nx, px), xp = bmap(x + 1, nx, px);
            const double den = (xm == x || xp == x) ? A.dx : 2.0 * A.dx;
            dc += (Bx[lin(g, xp, y, z)] - Bx[lin(g, xm, y, z)]) / den;
        }
        if (ny > 1) {
            const int ym = bmap(y - 1, ny, py), yp = bmap(y + 1, ny, py);
            const double den = (ym == y || yp == y) ? A.dy : 2.0 * A.dy;
            dc += (By[lin(g, x, yp, z)] - By[lin(g, x, ym, z)]) / den;
        }
        if (nz > 1) {
            const int zm = bmap(z - 1, nz, pz), zp = bmap(z + 1, nz, pz);
            const double den = (zm == z || zp == z) ? A.dz : 2.0 * A.dz;
            dc += (Bz[lin(g, x, y, zp)] - Bz[lin(g, x, y, zm)]) / den;
        }
        mcell = fmax(mcell, fabs(dc));
    }
    if (A.floors)
        for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < A.nfloor; k += (long long)gridDim.x * 256) {
            frho += (double)A.floors[2 * k];
            fp += (double)A.floors[2 * k + 1];
        }
    __shared__ double red4[4];
    double* out = A.part + (long long)blockIdx.x * DIAG_NV;
    double r;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        r = block_reduce<false>(sum[m], red4);
        if (threadIdx.x == 0) out[DG_SUM + m] = r;
    }
    r = block_reduce<true>(mct, red4);
    if (threadIdx.x == 0) out[DG_DIVCT] = r;
    r = block_reduce<true>(mcell, red4);
    if (threadIdx.x == 0) out[DG_DIVC] = r;
    r = block_reduce<true>(nrho, red4);
    if (threadIdx.x == 0) out[DG_NRHO] = r;
    r = block_reduce<true>(np, red4);
    if (threadIdx.x == 0) out[DG_NP] = r;
    r = block_reduce<false>(frho, red4);
    if (threadIdx.x == 0) out[DG_NFRHO] = r;
    r = block_reduce<false>(fp, red4);
    if (threadIdx.x == 0) out[DG_NFP] = r;
}

// Combines the nb block partials of k_diag into out[0..DIAG_NV-1] (one
// block; thread t takes partials t, t + 256, ... in order).
__global__ __launch_bounds__(256) void k_diag_fold(const double* part, int nb, double* out) {
    __shared__ double red4[4];
#pragma unroll 1
    for (int v = 0; v < DIAG_NV; ++v) {
        const bool mx = (v == DG_DIVCT || v == DG_DIVC || v == DG_NRHO || v == DG_NP);
        double a = mx ? -INFINITY : 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) {
            const double x = part[(long long)b * DIAG_NV + v];
            a = mx ? fmax(a, x) : a + x;
        }
        const double r = mx ? block_reduce<true>(a, red4) : block_reduce<false>(a, red4);
        if (threadIdx.x == 0) out[v] = r;
    }
}

// initialize_bfield_from_q, src/fluxct.cpp:65-107, on active faces
__global__ __launch_bounds__(256) void k_init_faces(const double* q, double* b, Geom g) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.ncell) return;
    int x, y, z;
    cell_of(g, i, x, y, z);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const double* Bx = q + 4 * g.ncell;
    const double* By = q + 5 * g.ncell;
    const double* Bz = q + 6 * g.ncell;
    const int px = g.per[0], py = g.per[1], pz = g.per[2];
    b[i] = (-Bx[lin(g, bmap(x - 1, nx, px), y, z)] + 9.0 * Bx[i] + 9.0 * Bx[lin(g, bmap(x + 1, nx, px), y, z)] -
            Bx[lin(g, bmap(x + 2, nx, px), y, z)]) / 16.0;
    b[g.ncell + i] = (-By[lin(g, x, bmap(y - 1, ny, py), z)] + 9.0 * By[i] + 9.0 * By[lin(g, x, bmap(y + 1, ny, py), z)] -
                      By[lin(g, x, bmap(y + 2, ny, py), z)]) / 16.0;
    b[2 * g.ncell + i] = (-Bz[lin(g, x, y, bmap(z - 1, nz, pz))] + 9.0 * Bz[i] + 9.0 * Bz[lin(g, x, y, bmap(z + 1, nz, pz))] -
                          Bz[lin(g, x, y, bmap(z + 2, nz, pz))]) / 16.0;
}

// ------------------------------------------- host layout <-> device layout
// One variable in the reference's host layout: (nx+6) x (ny+6) x (nz+6)
// values, three ghost layers per side, z fastest (the q(m, i, j, k) view of
// src/bound.cpp:19-95 for one m); on the device: active cells only, x
// fastest.  Both directions transpose (x, z) tiles of 32 x 32 through LDS at
// fixed y, so that global reads and writes are contiguous along a wavefront
// on both sides.  Block 32 x 8; grid (z tiles, x tiles, y planes), the
// planes strided by gridDim.z.
constexpr int LAY_T = 32;

// host -> device: active cells of the reference layout
__global__ __launch_bounds__(256) void k_ref_to_dev(const double* __restrict__ ref, double* __restrict__ dev, Geom g) {
    __shared__ double t[LAY_T][LAY_T + 1];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const size_t n6y = (size_t)ny + 6, n6z = (size_t)nz + 6;
    const int z0 = blockIdx.x * LAY_T, x0 = blockIdx.y * LAY_T;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int y = blockIdx.z; y < ny; y += gridDim.z) {
#pragma unroll
        for (int r = 0; r < LAY_T; r += 8) {
            const int x = x0 + ty + r, z = z0 + tx;
            if (x < nx && z < nz) t[ty + r][tx] = ref[(((size_t)x + 3) * n6y + (size_t)y + 3) * n6z + (size_t)z + 3];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < LAY_T; r += 8) {
            const int z = z0 + ty + r, x = x0 + tx;
            if (x < nx && z < nz) dev[((size_t)z * ny + y) * nx + x] = t[tx][ty + r];
        }
        __syncthreads();
    }
}

// device -> host: every cell of the reference layout, the ghost layers filled
// with the active cells the boundary conditions map them to (src/bound.cpp)
__global__ __launch_bounds__(256) void k_dev_to_ref(const double* __restrict__ dev, double* __restrict__ ref, Geom g) {
    __shared__ double t[LAY_T][LAY_T + 1];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int n6x = nx + 6, n6y = ny + 6, n6z = nz + 6;
    const int k0 = blockIdx.x * LAY_T, i0 = blockIdx.y * LAY_T;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int j = blockIdx.z; j < n6y; j += gridDim.z) {
        const size_t yo = (size_t)bmap(j - 3, ny, g.per[1]) * nx;
#pragma unroll
        for (int r = 0; r < LAY_T; r += 8) {
            const int k = k0 + ty + r, i = i0 + tx;
            if (i < n6x && k < n6z)
                t[ty + r][tx] = dev[(size_t)bmap(k - 3, nz, g.per[2]) * ny * nx + yo + bmap(i - 3, nx, g.per[0])];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < LAY_T; r += 8) {
            const int i = i0 + ty + r, k = k0 + tx;
            if (i < n6x && k < n6z) ref[((size_t)i * n6y + j) * n6z + k] = t[tx][ty + r];
        }
        __syncthreads();
    }
}

}  // namespace howmhd
