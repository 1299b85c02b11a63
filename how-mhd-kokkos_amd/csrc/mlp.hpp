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
#include <type_traits>

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

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Device form of the weights for the sweep's network: per-lane A fragments of
// the three products (MFMA 32x32x16 bf16, two k-steps each, see
// mlp_forward_wave), the FP32 biases of layers 1 and 2 in accumulator-row
// order, and the FP32 output bias.
struct MlpDev {
    alignas(16) uint16_t W1f[64][2][8];
    alignas(16) uint16_t W2f[64][2][8];
    alignas(16) uint16_t W3f[64][2][8];
    alignas(16) float b1r[2][16];
    alignas(16) float b2r[2][16];
    float b3;
    float logit_threshold;       // flag when z3 > threshold (0 <-> p > 0.5)
};

// Output unit held in accumulator row r (0..15) by the lanes of half-wave h
// (lanes 32h..32h+31) of a 32x32 MFMA result; the lane's column is the
// sample l % 32.
__host__ __device__ constexpr int mlp_acc_row(int r, int h) { return 8 * (r / 4) + 4 * h + r % 4; }
// Input unit of k-slot v (0..7) of k-step s seen by half-wave h in layers 2
// and 3, whose B operand is the previous layer's accumulator as it lies in
// the lane (rows 8s..8s+7).
__host__ __device__ constexpr int mlp_hidden_slot(int s, int v, int h) { return mlp_acc_row(8 * s + v, h); }

// B-operand fragments of one 32x32 bf16 matrix Wt[k][n]: [k-step s][n-tile u]
struct BFrag {
    bf16x4 f[2][2];
};

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
#if HOWMHD_EXACT
        const float rspan = 1.0f / span;
#else
        const float rspan = __builtin_amdgcn_rcpf(span);  // FMA forms: hardware reciprocal (<= 1 ulp)
#endif
        const float inv = span > MLP_FLAT * fmaxf(fabsf(hi), fabsf(lo)) + 1e-30f ? rspan : 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) f[6 * g + k] = (src[g][k] - lo) * inv;
    }
}

// The forward pass of the sweeps (routing and in-sweep logits): the 64
// lanes' interfaces are two 32-sample tiles, each layer is one pair of
// v_mfma_f32_32x32x16_bf16 products D = W X^T (output units x samples), and
// everything stays in registers.
//   * Layer-1 operand X^T: k-slot (h, v) of k-step 0 is feature 8h + v, of
//     k-step 1 feature 16 + v (h = 0, v < 2; the rest is 0).  Lane l holds
//     its own interface's features, so the half-waves trade halves with
//     v_permlane32_swap: one swap yields the operands of both tiles.
//   * A 32x32 result leaves output unit mlp_acc_row(r, h) of sample l % 32 in
//     row r of lane l.  The next layer takes these 16 rows, ReLU'd and
//     rounded to bf16, as its B operand unchanged (rows 0..7: k-step 0, rows
//     8..15: k-step 1); the weights' A fragments follow the same k order
//     (mlp_hidden_slot), so no data moves between layers.
//   * The output layer's A operand is w3 in row 0: z3 - b3 of sample l of the
//     tile lands in row 0 of lane l < 32, and one more swap returns each
//     tile's logits to the lanes of its interfaces.
// Per sample the products are the standalone inference kernel's (bf16
// operands, FP32 accumulation); only the order of the FP32 sums differs.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

// bf16(max(a, 0)), bf16(max(b, 0)) as one dword: rounding keeps the sign, so
// the ReLU is an integer max of the bf16 patterns (negative and -0 -> +0)
__device__ __forceinline__ uint32_t relu_bf2(float a, float b) {
    const s16x2 v = __builtin_bit_cast(s16x2, f2bf2(a, b));
    const s16x2 z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, z));
}

__device__ __forceinline__ f32x16 mfma32(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ void relu_pack(const f32x16& c, u32x4& lo, u32x4& hi) {
    lo = u32x4{relu_bf2(c[0], c[1]), relu_bf2(c[2], c[3]), relu_bf2(c[4], c[5]), relu_bf2(c[6], c[7])};
    hi = u32x4{relu_bf2(c[8], c[9]), relu_bf2(c[10], c[11]), relu_bf2(c[12], c[13]), relu_bf2(c[14], c[15])};
}

template <int B, int E, typename F>
__device__ __forceinline__ void sfor_mlp(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor_mlp<B + 1, E>(f);
    }
}

__device__ __forceinline__ float mlp_forward_wave(const float* feat, const MlpDev& m) {
    const int l = threadIdx.x & 63;
    const int h = l >> 5;
    uint32_t P[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) P[d] = f2bf2(feat[2 * d], feat[2 * d + 1]);
    u32x4 x0[2], x1[2];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const auto r = __builtin_amdgcn_permlane32_swap(P[d], P[4 + d], false, false);
        x0[0][d] = r[0];
        x0[1][d] = r[1];
    }
    {
        const auto r = __builtin_amdgcn_permlane32_swap(P[8], 0u, false, false);
        x1[0] = u32x4{r[0], 0u, 0u, 0u};
        x1[1] = u32x4{r[1], 0u, 0u, 0u};
    }
    const u32x4* w1 = reinterpret_cast<const u32x4*>(&m.W1f[l][0][0]);
    const u32x4* w2 = reinterpret_cast<const u32x4*>(&m.W2f[l][0][0]);
    const u32x4* w3 = reinterpret_cast<const u32x4*>(&m.W3f[l][0][0]);
    // The biases in accumulator-row order differ only between the half-waves.
    // Each lane loads one value per layer, row l % 16 of its half, and a DPP
    // row_newbcast:r copies lane r of every 16-lane row to the whole row, so
    // every lane of half h gets row r of half h: two 4-byte loads per lane
    // instead of eight 16-byte ones.
    const int vb1 = __float_as_int(m.b1r[h][l & 15]);
    const int vb2 = __float_as_int(m.b2r[h][l & 15]);
    f32x16 b1, b2;
    sfor_mlp<0, 16>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        b1[r] = __int_as_float(__builtin_amdgcn_update_dpp(0, vb1, 0x150 + r, 0xf, 0xf, false));
        b2[r] = __int_as_float(__builtin_amdgcn_update_dpp(0, vb2, 0x150 + r, 0xf, 0xf, false));
    });
    const u32x4 w1a = w1[0], w1b = w1[1], w2a = w2[0], w2b = w2[1], w3a = w3[0], w3b = w3[1];
    float z[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        f32x16 c = mfma32(w1a, x0[t], b1);
        c = mfma32(w1b, x1[t], c);
        u32x4 lo, hi;
        relu_pack(c, lo, hi);
        c = mfma32(w2a, lo, b2);
        c = mfma32(w2b, hi, c);
        relu_pack(c, lo, hi);
        f32x16 o = {};
        o = mfma32(w3a, lo, o);
        o = mfma32(w3b, hi, o);
        z[t] = o[0];
    }
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(z[0]), __float_as_uint(z[1]), false, false);
    return __uint_as_float(r[0]) + m.b3;
}

}  // namespace howmhd
