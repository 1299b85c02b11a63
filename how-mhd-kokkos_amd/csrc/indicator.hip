// Learned smoothness indicator: inference, training (forward + backward with
// the weight gradients on bf16 MFMA, FP32 accumulation) and Adam, behind the
// C-ABI of include/howmhd_mlp.h.
//
// Training numerics (the CPU check in tests/ follows the same definitions):
//   forward  x, W1, W2 rounded to bf16; z1 = x W1^T + b1 (FP32 accumulate);
//            a1 = bf16(relu(z1)); z2 = a1 W2^T + b2; a2 = bf16(relu(z2));
//            z3 = b3 + sum_j a2_j * bf16(w3_j) in FP32, j = 0..31
//   loss     w * (softplus(z3) - y z3), w = pos_weight for y = 1 else 1
//   backward dz3 = w (sigmoid(z3) - y) * scale (FP32)
//            dw3 = sum_i dz3 a2, db3 = sum_i dz3 (FP32)
//            dz2 = bf16(dz3 * bf16(w3) where a2 != 0); dW2 = dz2^T a1; db2 = sum_i dz2
//            dz1 = bf16((dz2 W2) where a1 != 0);       dW1 = dz1^T x;  db1 = sum_i dz1
#include "howmhd_mlp.h"
#include "mlp.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

using namespace howmhd;

namespace {

constexpr uint16_t BF16_ONE = 0x3F80;

// B fragments and vectors of the network, rebuilt by each block from the FP32
// master parameters.
struct NetFrags {
    BFrag W1f;   // B[k][n] = W1[n][k] (k < 18)
    BFrag W2f;   // B[k][n] = W2[n][k]
    BFrag W2b;   // B[n][k] = W2[n][k]   (backward: dz2 * W2)
};

__device__ __forceinline__ void build_frags(const float* P, NetFrags& F) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kk = 16 * s + 4 * (l / 16) + v;  // fragment row
                const int cc = 16 * u + (l % 16);          // fragment column
                F.W1f.f[s][u][v] = (short)(kk < MLP_IN ? f2bf(P[MLP_OFF_W1 + cc * MLP_IN + kk]) : 0);
                F.W2f.f[s][u][v] = (short)f2bf(P[MLP_OFF_W2 + cc * MLP_H + kk]);
                F.W2b.f[s][u][v] = (short)f2bf(P[MLP_OFF_W2 + kk * MLP_H + cc]);
            }
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// sum over the 64 rows (i) of two bf16 row buffers: D[n][k] = sum_i A[i][n] B[i][k]
// for n-tile m and k-tile u (one 16 x 16 tile per call)
__device__ __forceinline__ f32x4 gemm_tn_tile(const uint16_t* A, const uint16_t* B, int m, int u, bool b_ones) {
    const int l = threadIdx.x & 63;
    const int r16 = l % 16, kq = 4 * (l / 16);
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        bf16x4 a, b;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * s + kq + v;
            a[v] = (short)A[i * MLP_ROW + 16 * m + r16];
            b[v] = (short)(b_ones ? BF16_ONE : B[i * MLP_ROW + 16 * u + r16]);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    return c;
}

constexpr int WAVES = 4;

__global__ __launch_bounds__(256) void k_mlp_infer(const float* X, int64_t n, const float* P, float* logits) {
    __shared__ uint16_t act[WAVES][64 * MLP_ROW];
    __shared__ float vec[3 * MLP_H + 1];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int j = threadIdx.x; j < 3 * MLP_H + 1; j += 256) {
        float v;
        if (j < MLP_H) v = P[MLP_OFF_B1 + j];
        else if (j < 2 * MLP_H) v = P[MLP_OFF_B2 + j - MLP_H];
        else if (j < 3 * MLP_H) v = bf2f(f2bf(P[MLP_OFF_W3 + j - 2 * MLP_H]));
        else v = P[MLP_OFF_B3];
        vec[j] = v;
    }
    __syncthreads();
    NetFrags F;
    build_frags(P, F);
    const int64_t i = ((int64_t)blockIdx.x * WAVES + w) * 64 + l;
    const bool valid = i < n;
    uint16_t* a = act[w];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) a[l * MLP_ROW + j] = (valid && j < MLP_IN) ? f2bf(X[i * MLP_IN + j]) : 0;
    __builtin_amdgcn_wave_barrier();
    layer_relu(a, a, F.W1f, vec);
    layer_relu(a, a, F.W2f, vec + MLP_H);
    float z = vec[3 * MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z += bf2f(a[l * MLP_ROW + j]) * vec[2 * MLP_H + j];
    if (valid) logits[i] = z;
}

__global__ __launch_bounds__(256) void k_mlp_grad(const float* X, const float* Y, int64_t n, const float* P,
                                                  float pos_weight, float scale, float* G, float* loss) {
    __shared__ uint16_t xs[WAVES][64 * MLP_ROW], a1s[WAVES][64 * MLP_ROW], a2s[WAVES][64 * MLP_ROW];
    __shared__ uint16_t d2s[WAVES][64 * MLP_ROW], d1s[WAVES][64 * MLP_ROW];
    __shared__ float vec[3 * MLP_H + 1];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int r16 = l % 16, kq = 4 * (l / 16);
    for (int j = threadIdx.x; j < 3 * MLP_H + 1; j += 256) {
        float v;
        if (j < MLP_H) v = P[MLP_OFF_B1 + j];
        else if (j < 2 * MLP_H) v = P[MLP_OFF_B2 + j - MLP_H];
        else if (j < 3 * MLP_H) v = bf2f(f2bf(P[MLP_OFF_W3 + j - 2 * MLP_H]));
        else v = P[MLP_OFF_B3];
        vec[j] = v;
    }
    __syncthreads();
    NetFrags F;
    build_frags(P, F);
    const int64_t i = ((int64_t)blockIdx.x * WAVES + w) * 64 + l;
    const bool valid = i < n;
    uint16_t* x = xs[w];
    uint16_t* a1 = a1s[w];
    uint16_t* a2 = a2s[w];
    uint16_t* d2 = d2s[w];
    uint16_t* d1 = d1s[w];

    // ---- forward
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) x[l * MLP_ROW + j] = (valid && j < MLP_IN) ? f2bf(X[i * MLP_IN + j]) : 0;
    __builtin_amdgcn_wave_barrier();
    layer_relu(x, a1, F.W1f, vec);
    layer_relu(a1, a2, F.W2f, vec + MLP_H);
    const float* w3 = vec + 2 * MLP_H;
    float z = vec[3 * MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z += bf2f(a2[l * MLP_ROW + j]) * w3[j];
    const float yv = valid ? Y[i] : 0.f;
    const float wgt = (yv > 0.5f) ? pos_weight : 1.f;
    const float p = 1.f / (1.f + expf(-z));
    const float li = valid ? wgt * (fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))) - yv * z) : 0.f;
    const float dz3 = valid ? wgt * (p - yv) * scale : 0.f;

    // ---- output layer
    const float lsum = wave_sum(li * scale);
    const float db3 = wave_sum(dz3);
    if (l == 0) {
        atomicAdd(loss, lsum);
        atomicAdd(&G[MLP_OFF_B3], db3);
    }
#pragma unroll 4
    for (int j = 0; j < MLP_H; ++j) {
        const float a = bf2f(a2[l * MLP_ROW + j]);
        const float g = wave_sum(dz3 * a);
        if (l == 0) atomicAdd(&G[MLP_OFF_W3 + j], g);
        d2[l * MLP_ROW + j] = (a != 0.f) ? f2bf(dz3 * w3[j]) : 0;
    }
    __builtin_amdgcn_wave_barrier();

    // ---- layer 2: dW2 = dz2^T a1, db2 = sum_i dz2
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const f32x4 c = gemm_tn_tile(d2, a1, m, u, false);
#pragma unroll
            for (int v = 0; v < 4; ++v) atomicAdd(&G[MLP_OFF_W2 + (16 * m + kq + v) * MLP_H + 16 * u + r16], c[v]);
        }
        const f32x4 c = gemm_tn_tile(d2, nullptr, m, 0, true);
        if (r16 == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) atomicAdd(&G[MLP_OFF_B2 + 16 * m + kq + v], c[v]);
        }
    }
    // ---- dz1 = (dz2 W2) where a1 != 0
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        bf16x4 a[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[s][v] = (short)d2[(16 * t + r16) * MLP_ROW + 16 * s + kq + v];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], F.W2b.f[0][u], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], F.W2b.f[1][u], c, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * t + kq + v, col = 16 * u + r16;
                d1[row * MLP_ROW + col] = (a1[row * MLP_ROW + col] != 0) ? f2bf(c[v]) : 0;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    // ---- layer 1: dW1 = dz1^T x (k < 18), db1 = sum_i dz1
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const f32x4 c = gemm_tn_tile(d1, x, m, u, false);
            const int k = 16 * u + r16;
            if (k < MLP_IN) {
#pragma unroll
                for (int v = 0; v < 4; ++v) atomicAdd(&G[MLP_OFF_W1 + (16 * m + kq + v) * MLP_IN + k], c[v]);
            }
        }
        const f32x4 c = gemm_tn_tile(d1, nullptr, m, 0, true);
        if (r16 == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) atomicAdd(&G[MLP_OFF_B1 + 16 * m + kq + v], c[v]);
        }
    }
}

__global__ void k_mlp_adam(float* P, float* M, float* V, const float* G, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= MLP_NPARAM) return;
    const float g = G[j];
    const float m = b1 * M[j] + (1.f - b1) * g;
    const float v = b2 * V[j] + (1.f - b2) * g * g;
    M[j] = m;
    V[j] = v;
    P[j] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

extern "C" {

int howmhd_mlp_infer(const float* X, int64_t n, const float* params, float* logits, void* stream) {
    if (n <= 0) return HOWMHD_OK;
    if (!X || !params || !logits) return HOWMHD_EINVAL;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_infer, dim3(nb), dim3(256), 0, (hipStream_t)stream, X, n, params, logits);
    return hipGetLastError() == hipSuccess ? HOWMHD_OK : HOWMHD_EHIP;
}

int howmhd_mlp_grad(const float* X, const float* y, int64_t n, const float* params, float pos_weight,
                    float scale, float* grad, float* loss, void* stream) {
    if (n <= 0) return HOWMHD_OK;
    if (!X || !y || !params || !grad || !loss) return HOWMHD_EINVAL;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, (hipStream_t)stream, X, y, n, params, pos_weight, scale,
                       grad, loss);
    return hipGetLastError() == hipSuccess ? HOWMHD_OK : HOWMHD_EHIP;
}

int howmhd_mlp_adam(float* params, float* m, float* v, const float* grad, int64_t t, float lr, float beta1,
                    float beta2, float eps, void* stream) {
    if (!params || !m || !v || !grad || t < 1) return HOWMHD_EINVAL;
    const float c1 = 1.f - powf(beta1, (float)t);
    const float c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, m, v,
                       grad, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? HOWMHD_OK : HOWMHD_EHIP;
}

}  // extern "C"
