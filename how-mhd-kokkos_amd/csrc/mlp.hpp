// Learned smoothness indicator: an 18 -> 32 -> 32 -> 1 MLP (ReLU, ReLU,
// sigmoid) evaluated with bf16 MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulation) by whole wavefronts: the 64 lanes of a wavefront are 64
// interfaces (or training samples), split into four 16-row tiles.
//
// Inputs: the 6-cell stencils (the WENO5 stencil of the interface) of rho, p
// and |B|, each mapped to [0,1] by its own min and max over the 6 cells.
// Activations between layers are rounded to bf16 (they are MFMA operands);
// the last layer is an FP32 dot product with the bf16 output weights.
//
// MFMA 16x16x16 operand layout (lane l, element v = 0..3):
//   A (16 x 16, rows x k):  A[l % 16][4*(l/16) + v]
//   B (16 x 16, k x cols):  B[4*(l/16) + v][l % 16]
//   D (16 x 16, rows x cols): D[4*(l/16) + v][l % 16]
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace howmhd {

constexpr int MLP_IN = 18, MLP_H = 32;
// master (FP32) parameter vector: W1[32][18], b1[32], W2[32][32], b2[32], w3[32], b3
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 576, MLP_OFF_W2 = 608, MLP_OFF_B2 = 1632, MLP_OFF_W3 = 1664,
              MLP_OFF_B3 = 1696, MLP_NPARAM = 1697;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even: gfx950's v_cvt_pk_bf16_f32
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// two values in one conversion: low half a, high half b
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// four values as one bf16x4 (64-bit LDS store)
__device__ __forceinline__ bf16x4 f2bf4(float a, float b, float c, float d) {
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 w = {f2bf2(a, b), f2bf2(c, d)};
    return __builtin_bit_cast(bf16x4, w);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Device form of the weights for inference: bf16 B operands of layers 1 and
// 2 (transposed, layer-1 k padded 18 -> 32 with zeros), FP32 biases, bf16
// output weights held as floats.
struct MlpDev {
    uint16_t W1t[MLP_H][MLP_H];  // W1t[k][n] = W1[n][k], k < 18
    uint16_t W2t[MLP_H][MLP_H];  // W2t[k][n] = W2[n][k]
    float b1[MLP_H], b2[MLP_H], w3[MLP_H];
    float b3;
    float logit_threshold;       // flag when z3 > threshold (0 <-> p > 0.5)
};

// B-operand fragments of one 32x32 bf16 matrix Wt[k][n]: [k-step s][n-tile u]
struct BFrag {
    bf16x4 f[2][2];
};

__device__ __forceinline__ void load_bfrag(const uint16_t (*Wt)[MLP_H], BFrag& b) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) b.f[s][u][v] = (short)Wt[16 * s + 4 * (l / 16) + v][16 * u + (l % 16)];
}

// Scratch rows of one wavefront: 64 rows x 32 bf16 (+2 pad to stagger banks).
constexpr int MLP_ROW = MLP_H + 2;

// One dense layer on the 64 rows of a wavefront: dst <- relu(src * Wt + b),
// bf16 rows of MLP_ROW, four 16-row tiles.  src may equal dst: every lane
// reads its fragments of a tile before any lane writes that tile.
__device__ __forceinline__ void layer_relu(const uint16_t* src, uint16_t* dst, const BFrag& W, const float* bias) {
    const int l = threadIdx.x & 63;
    const int r16 = l % 16, kq = 4 * (l / 16);
    const float bl0 = bias[r16], bl1 = bias[16 + r16];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        bf16x4 a[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[s][v] = (short)src[(16 * t + r16) * MLP_ROW + 16 * s + kq + v];
        f32x4 d[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const float b = u ? bl1 : bl0;
            f32x4 c = {b, b, b, b};
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], W.f[0][u], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], W.f[1][u], c, 0, 0, 0);
            d[u] = c;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = d[u][v];
                dst[(16 * t + kq + v) * MLP_ROW + 16 * u + r16] = f2bf(z > 0.f ? z : 0.f);
            }
        __builtin_amdgcn_wave_barrier();
    }
}

// Features of one interface from its 6-cell stencils of rho, p and |B|: each
// mapped to [0, 1] by its min and max, or all 0 when max - min <= MLP_FLAT
// max(|min|, |max|).  The relative floor keeps round-off-level wiggles of
// flat data (which leave the WENO weights ideal) from being stretched into
// 0/1 patterns that look like jumps.
constexpr float MLP_FLAT = 1e-5f;

__device__ __forceinline__ void mlp_features(const float* rho, const float* p, const float* b, float* f) {
    const float* src[3] = {rho, p, b};
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        float lo = src[g][0], hi = src[g][0];
#pragma unroll
        for (int k = 1; k < 6; ++k) {
            lo = fminf(lo, src[g][k]);
            hi = fmaxf(hi, src[g][k]);
        }
        const float span = hi - lo;
        const float inv = span > MLP_FLAT * fmaxf(fabsf(hi), fabsf(lo)) + 1e-30f ? 1.0f / span : 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) f[6 * g + k] = (src[g][k] - lo) * inv;
    }
}

// The forward pass the routed sweep uses.  Transposed products: each layer is
// D = W A^T, with the weights as the MFMA A operand (the B fragments of
// load_bfrag hold exactly W[16u + l%16][16s + 4(l/16) + v]) and the
// activations of 16 samples as the B operand, so that every lane reads and
// writes its activations as 4 consecutive bf16 (64-bit LDS accesses); the
// output layer is a third product whose A operand is w3 in row 0, and the
// logits are moved to their lanes with a cross-lane permute.
// act: 16 x MLP_ROWT bf16 per wavefront (8-byte aligned rows).
constexpr int MLP_ROWT = 36;

__device__ __forceinline__ bf16x4 lds_b64(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ __forceinline__ void layer_relu_t(uint16_t* act, const BFrag& W, const float* bias) {
    const int l = threadIdx.x & 63;
    const int r16 = l % 16, kq = 4 * (l / 16);
    const bf16x4 b0 = lds_b64(act + r16 * MLP_ROWT + kq);
    const bf16x4 b1 = lds_b64(act + r16 * MLP_ROWT + 16 + kq);
    f32x4 d[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        f32x4 c = {bias[16 * u + kq], bias[16 * u + kq + 1], bias[16 * u + kq + 2], bias[16 * u + kq + 3]};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W.f[0][u], b0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W.f[1][u], b1, c, 0, 0, 0);
        d[u] = c;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 2; ++u)
        *reinterpret_cast<bf16x4*>(act + r16 * MLP_ROWT + 16 * u + kq) =
            f2bf4(fmaxf(d[u][0], 0.f), fmaxf(d[u][1], 0.f), fmaxf(d[u][2], 0.f), fmaxf(d[u][3], 0.f));
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float mlp_forward_wave_t(const float* feat, uint16_t* act, const BFrag& W1,
                                                    const BFrag& W2, const MlpDev& m) {
    const int l = threadIdx.x & 63;
    const int grp = l >> 4, r16 = l & 15, kq = 4 * (l >> 4);
    // w3 as the A operand: row 0 only
    bf16x4 w3f[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) w3f[s][v] = (short)(r16 == 0 ? f2bf(m.w3[16 * s + kq + v]) : 0);
    float zl = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (grp == t) {
#pragma unroll
            for (int j = 0; j < MLP_H; j += 4) {
                auto fv = [&](int k) { return k < MLP_IN ? feat[k < MLP_IN ? k : 0] : 0.f; };
                *reinterpret_cast<bf16x4*>(act + r16 * MLP_ROWT + j) = f2bf4(fv(j), fv(j + 1), fv(j + 2), fv(j + 3));
            }
        }
        __builtin_amdgcn_wave_barrier();
        layer_relu_t(act, W1, m.b1);
        layer_relu_t(act, W2, m.b2);
        const bf16x4 b0 = lds_b64(act + r16 * MLP_ROWT + kq);
        const bf16x4 b1 = lds_b64(act + r16 * MLP_ROWT + 16 + kq);
        f32x4 c = {kq == 0 ? m.b3 : 0.f, 0.f, 0.f, 0.f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3f[0], b0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3f[1], b1, c, 0, 0, 0);
        // lanes 0..15 hold z3 of samples 0..15 of this tile in c[0]
        const float z = __shfl(c[0], r16, 64);
        if (grp == t) zl = z;
        __builtin_amdgcn_wave_barrier();
    }
    return zl;
}

}  // namespace howmhd
