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

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Adds the per-wavefront counters of the routed sweeps (n slots of
// (flagged, total, wavefronts flagged, wavefronts total)) to the four totals
// and clears the slots.  Grid-stride loop over the slots; one atomic per
// total and block.
__global__ __launch_bounds__(256) void k_fold_wave_stats(uint4* w, long long n, unsigned long long* tot) {
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
        atomicAdd(&tot[k], red[k][0] + red[k][1] + red[k][2] + red[k][3]);
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
