// Kernels of the per-step calls (tstep and the face initialisation), built
// once, in the plain FP64 unit (-ffp-contract=off): the CFL maximum and dt
// are the reference's bit for bit for every form.
#pragma once

#include "kernels.hpp"

namespace howmhd {

// ---------------------------------------------------------------- CFL (tstep)
// Per-direction maximum of |v_d| + c_fast,d over active cells of q(0)
// (src/tstep.cpp:28-83).  Non-negative doubles order like their bit patterns,
// so the block maxima are combined with a 64-bit atomicMax.
__global__ __launch_bounds__(256) void k_cfl(const double* q, Geom g, Phys ph, unsigned long long* out) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    const double gam = ph.gam;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < g.ncell; i += (long long)gridDim.x * 256) {
        const double DD = q[i], Mx = q[g.ncell + i], My = q[2 * g.ncell + i], Mz = q[3 * g.ncell + i];
        const double Bx = q[4 * g.ncell + i], By = q[5 * g.ncell + i], Bz = q[6 * g.ncell + i], EE = q[7 * g.ncell + i];
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
    // wavefront reduction (64 lanes), then across the 4 wavefronts
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, off, 64));
        m1 = fmax(m1, __shfl_xor(m1, off, 64));
        m2 = fmax(m2, __shfl_xor(m2, off, 64));
    }
    __shared__ double red[3][4];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { red[0][w] = m0; red[1][w] = m1; red[2][w] = m2; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        const double r = fmax(fmax(red[d][0], red[d][1]), fmax(red[d][2], red[d][3]));
        atomicMax(&out[d], (unsigned long long)__double_as_longlong(r));
    }
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

}  // namespace howmhd
