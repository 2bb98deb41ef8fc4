// xts_mlp.hpp -- the surrogate network of the far-pixel cubature: a
// 4 -> 64 -> 64 -> 64 -> 1 multilayer perceptron with SiLU activations, its
// parameter layout, and its BF16 inference on the matrix cores.
//
// What it predicts.  For a pixel whose rule points all lie within two
// consecutive intervals of the PSF table's separation grid (one node, at
// separation s1, inside the points' range), the cubature reduces to three
// row weights, which are linear in three sums over the rule points: the
// weight total, A = sum w sigma and B1 = sum w |sigma|, with sigma = (s -
// s1) / (s1 - s0) the point's separation in units of the interval below the
// node.  A is smooth in the pixel's geometry and a 2x2 Gauss-Legendre rule
// of sigma gives it; B1 is not (it has a kink wherever a point crosses the
// node), and is what the network predicts, as the excess y = B1 - |A| >= 0,
// from four features of sigma over the pixel's bilinear parameter square
// (u, v in [-1, 1]): its centre value, its half-differences across the
// pixel along u and along v, and the sum of its four edge-midpoint second
// differences (the curvature, which for a pixel far from the source lies
// across the radial direction).  B1 is energy independent, so one network
// evaluation serves every energy plane; the pixel-integrated PSF of each
// energy follows exactly from the three row weights and the table.
//
// Parameters (FP32, 8705): W1 [64][4], b1 [64], W2 [64][64], b2 [64],
// W3 [64][64], b3 [64], w4 [64], b4; plus a normalisation record
// (mu [4], inv_sd [4], y_scale, y_offset): x' = (x - mu) inv_sd, y = y'
// y_scale + y_offset.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

namespace xts {

constexpr int kMlpIn = 4, kMlpH = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH, kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpH * kMlpH;
constexpr int kOffW4 = kOffB3 + kMlpH, kOffB4 = kOffW4 + kMlpH;
constexpr int kMlpParams = kOffB4 + 1;   // 8705
constexpr int kMlpNorm = 10;             // mu[4], inv_sd[4], y_scale, y_offset

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef short mlp_s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t mlp_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float mlp_from_bf16(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// The inference records hold the network in base-2 units: the first
// layer's weights and every layer's biases times -log2(e), the output
// weights times -ln(2).  A layer's accumulator then holds z = -log2(e) x for
// the trained pre-activation x, and mlp_act(z) = z / (1 + 2^z) is
// -log2(e) silu(x), because silu(x) = x / (1 + e^-x) and e^-x = 2^z.  The
// next layer's products with the unchanged hidden weights, plus its scaled
// bias, are again that layer's z (the factors -ln(2) and -log2(e) cancel),
// and the output's weights take the last -ln(2).  This saves the multiply
// by -log2(e) per activation.  The hardware base-2 exponential and the
// reciprocal are about 1 ulp each; hidden activations are rounded to BF16
// (or FP8) anyway.  For large |z|, 2^z goes to 0 (act = z) or to +inf
// (act = 0), as silu does.
constexpr float kMlpNegLog2e = -1.4426950408889634f;
constexpr float kMlpNegLn2 = -0.6931471805599453f;
__device__ inline float mlp_act(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
}

// Inference record in the matrix cores' operand order.  A 16x16 weight tile
// (outputs 16 To .., inputs 16 Ti ..) is 64 lanes of four BF16 values: lane
// l holds W[16 To + l % 16][16 Ti + 4 (l / 16) + j], j = 0..3, the A operand
// of v_mfma_f32_16x16x16_bf16 for a product W . H^T whose B operand, lane l,
// is sample l % 16's inputs 16 Ti + 4 (l / 16) + j -- exactly how the
// previous layer's result D (lane l: outputs 16 To + 4 (l / 16) + i of
// sample l % 16) lies in registers, so layers chain without a transpose.
// The first layer's 16-wide K holds the inputs' high and low BF16 parts
// against the weights' high and low parts (x_hi w_hi + x_lo w_hi + x_hi
// w_lo).  The hidden layers run on v_mfma_f32_16x16x32_bf16, a high and a
// low weight tile each: a 32-wide input chunk Ti is two of the previous
// layer's 16-output tiles side by side, lane l carrying in its eight k-slots
// inputs 32 Ti + 4 (l / 16) + j and 32 Ti + 16 + 4 (l / 16) + j, j = 0..3 --
// the A operand's slots hold the weights of the same inputs, so the pair
// multiplies as W . H^T without moving any activation between lanes.
__host__ __device__ inline int mlp_kslot(int Ti, int g, int j) {
    return 32 * Ti + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}
// P = 2: the hidden layers' weights as BF16 high and low parts (MlpPacked,
// FP32-level weights, 35.9 KB); P = 1: the high parts alone (MlpPackedH,
// plain BF16 weights, 19.5 KB: half the hidden layers' matrix instructions
// and LDS reads).
template <int P>
struct MlpPackedT {
    static constexpr int kParts = P;
    mlp_bf16x4 w1[4][64];          // [To][lane]
    mlp_s16x8 w2[P][8][64];        // [hi/lo][To * 2 + Ti][lane]
    mlp_s16x8 w3[P][8][64];
    float b1[kMlpH], b2[kMlpH], b3[kMlpH], w4[kMlpH];
    float b4, norm[kMlpNorm];
    float pad[5];
};
using MlpPacked = MlpPackedT<2>;
using MlpPackedH = MlpPackedT<1>;
static_assert(sizeof(MlpPacked) % 16 == 0 && sizeof(MlpPackedH) % 16 == 0, "MlpPacked size");

// Pack FP32 parameters into the inference record (one block of 256 threads),
// in the base-2 units of mlp_act.
template <int P>
__device__ inline void mlp_pack(const float* __restrict__ p, const float* __restrict__ norm,
                                MlpPackedT<P>* __restrict__ out, int tid, int nthr) {
    for (int k = tid; k < 4 * 64; k += nthr) {
        const int To = k / 64, l = k % 64, g = l / 16, o = 16 * To + l % 16;
        mlp_bf16x4 v;
        for (int j = 0; j < 4; ++j) {
            const float w = kMlpNegLog2e * p[kOffW1 + o * kMlpIn + j];
            const uint16_t hi = mlp_bf16(w), lo = mlp_bf16(w - mlp_from_bf16(hi));
            v[j] = (short)(g <= 1 ? hi : (g == 2 ? lo : 0));
        }
        out->w1[To][l] = v;
    }
    for (int k = tid; k < 2 * 8 * 64; k += nthr) {
        const int layer = k / (8 * 64), r = k % (8 * 64);
        const int t = r / 64, l = r % 64, To = t / 2, Ti = t % 2;
        const float* W = p + (layer == 0 ? kOffW2 : kOffW3);
        mlp_s16x8 hi8, lo8;
        for (int j = 0; j < 8; ++j) {
            const float w = W[(16 * To + l % 16) * kMlpH + mlp_kslot(Ti, l / 16, j)];
            const uint16_t hi = mlp_bf16(w);
            hi8[j] = (short)hi;
            lo8[j] = (short)mlp_bf16(w - mlp_from_bf16(hi));
        }
        mlp_s16x8(*dst)[8][64] = layer == 0 ? out->w2 : out->w3;
        dst[0][t][l] = hi8;
        if constexpr (P == 2) dst[1][t][l] = lo8;
    }
    for (int k = tid; k < kMlpH; k += nthr) {
        out->b1[k] = kMlpNegLog2e * p[kOffB1 + k];
        out->b2[k] = kMlpNegLog2e * p[kOffB2 + k];
        out->b3[k] = kMlpNegLog2e * p[kOffB3 + k];
        out->w4[k] = kMlpNegLn2 * p[kOffW4 + k];
    }
    if (tid == 0) {
        out->b4 = p[kOffB4];
        for (int k = 0; k < kMlpNorm; ++k) out->norm[k] = norm[k];
        for (int k = 0; k < 5; ++k) out->pad[k] = 0.0f;
    }
}

typedef float mlp_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned mlp_u32x2 __attribute__((ext_vector_type(2)));

// Four floats to BF16 by the hardware's packed conversion (v_cvt_pk_bf16_f32,
// round to nearest even, as mlp_bf16).
__device__ inline mlp_bf16x4 mlp_to_bf16x4(float a, float b, float c, float d) {
    const mlp_bf16x2 lo = __builtin_convertvector((mlp_f32x2{a, b}), mlp_bf16x2);
    const mlp_bf16x2 hi = __builtin_convertvector((mlp_f32x2{c, d}), mlp_bf16x2);
    const mlp_u32x2 u = {__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)};
    return __builtin_bit_cast(mlp_bf16x4, u);
}

__device__ inline mlp_bf16x4 mlp_act_bf16(mlp_f32x4& acc) {
    return mlp_to_bf16x4(mlp_act(acc[0]), mlp_act(acc[1]), mlp_act(acc[2]), mlp_act(acc[3]));
}

// Two of the previous layer's 16-output tiles as one 32-wide input chunk.
__device__ inline mlp_bf16x8 mlp_chunk(const mlp_bf16x4& a, const mlp_bf16x4& b) {
    const mlp_s16x8 v = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
    return __builtin_bit_cast(mlp_bf16x8, v);
}

__device__ inline mlp_f32x4 mlp_mfma32(const mlp_s16x8& w, const mlp_bf16x8& h,
                                       const mlp_f32x4& acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16x8, w), h, acc,
                                                   0, 0, 0);
}

// The normalised inputs of layer 1, by lane group g: the high BF16 parts
// (g = 0, 2), the low parts (g = 1), zero (g = 3).
__device__ inline mlp_bf16x4 mlp_input_parts(const float* norm, const float (&x)[4], int g) {
    float xn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xn[j] = (x[j] - norm[j]) * norm[4 + j];
    const mlp_bf16x4 hi = mlp_to_bf16x4(xn[0], xn[1], xn[2], xn[3]);
    if (g != 1) return g == 3 ? mlp_bf16x4{0, 0, 0, 0} : hi;
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = xn[j] - mlp_from_bf16((uint16_t)hi[j]);
    return mlp_to_bf16x4(r[0], r[1], r[2], r[3]);
}

// One wave evaluates 16 samples: lane l carries sample l % 16's raw
// features x[4] (every lane of the sample the same values); the result, for
// sample l % 16, is returned in every lane.  m: the record (LDS or global).
template <int P>
__device__ inline float mlp_infer16(const MlpPackedT<P>* __restrict__ m, const float (&x)[4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    // normalised inputs, split into BF16 high and low parts by lane group
    mlp_bf16x4 in = mlp_input_parts(m->norm, x, g);
    mlp_bf16x4 H[4];
#pragma unroll
    for (int To = 0; To < 4; ++To) {
        mlp_f32x4 acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = m->b1[16 * To + 4 * g + i];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m->w1[To][lane], in, acc, 0, 0, 0);
        H[To] = mlp_act_bf16(acc);
    }
    // hidden layer 2
    mlp_bf16x8 X[2] = {mlp_chunk(H[0], H[1]), mlp_chunk(H[2], H[3])};
    mlp_f32x4 acc2[4];
#pragma unroll
    for (int To = 0; To < 4; ++To) {
#pragma unroll
        for (int i = 0; i < 4; ++i) acc2[To][i] = m->b2[16 * To + 4 * g + i];
#pragma unroll
        for (int Ti = 0; Ti < 2; ++Ti) {
            if constexpr (P == 2)
                acc2[To] = mlp_mfma32(m->w2[1][To * 2 + Ti][lane], X[Ti], acc2[To]);
            acc2[To] = mlp_mfma32(m->w2[0][To * 2 + Ti][lane], X[Ti], acc2[To]);
        }
    }
#pragma unroll
    for (int To = 0; To < 4; ++To) H[To] = mlp_act_bf16(acc2[To]);
    X[0] = mlp_chunk(H[0], H[1]);
    X[1] = mlp_chunk(H[2], H[3]);
    // hidden layer 3, its activations kept in FP32 for the output dot product
    float y = 0.0f;
#pragma unroll
    for (int To = 0; To < 4; ++To) {
        mlp_f32x4 acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = m->b3[16 * To + 4 * g + i];
#pragma unroll
        for (int Ti = 0; Ti < 2; ++Ti) {
            if constexpr (P == 2) acc = mlp_mfma32(m->w3[1][To * 2 + Ti][lane], X[Ti], acc);
            acc = mlp_mfma32(m->w3[0][To * 2 + Ti][lane], X[Ti], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) y = fmaf(m->w4[16 * To + 4 * g + i], mlp_act(acc[i]), y);
    }
    // the four lane groups hold disjoint outputs of the same sample
    y += __shfl_xor(y, 16);
    y += __shfl_xor(y, 32);
    return fmaf(y + m->b4, m->norm[8], m->norm[9]);
}

// ---------------------------------------------------------------- FP8
// The FP8 (OCP e4m3) inference record: layer 1 as in MlpPacked (BF16 high
// and low parts: its four inputs carry the features' full precision), the
// two hidden layers' weights as e4m3 with one scale per layer (w = q sw,
// the layer's largest |w| at kFp8Max), in v_mfma_f32_16x16x32_fp8_fp8
// operand order: tile To * 2 + Ti, lane l's eight bytes are the weights of
// output 16 To + l % 16 for inputs mlp_kslot(Ti, l / 16, j), j = 0..7 (the
// contraction pairs A's and B's bytes by lane group and byte, so the
// previous layer's accumulator lanes feed B as in the BF16 record).
// Activations are quantized per wave and layer at run time: the wave's
// largest |h| at kFp8Max.  e4m3 has 3 mantissa bits (relative rounding
// 2^-4) and no saturation (beyond 464 it converts to NaN): kFp8Max = 440
// keeps every scaled value finite.
constexpr float kFp8Max = 440.0f;
typedef long mlp_fp8x8;
struct MlpPacked8 {
    mlp_bf16x4 w1[4][64];
    mlp_fp8x8 w2[8][64];
    mlp_fp8x8 w3[8][64];
    float b1[kMlpH], b2[kMlpH], b3[kMlpH], w4[kMlpH];
    float b4, norm[kMlpNorm];
    float sw2, sw3;
    float pad[3];
};
static_assert(sizeof(MlpPacked8) % 16 == 0, "MlpPacked8 size");

// Eight floats to e4m3 bytes (v_cvt_pk_fp8_f32, round to nearest even).
__device__ inline mlp_fp8x8 mlp_fp8_pack8(const float (&v)[8]) {
    int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
    int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
    return (mlp_fp8x8)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Pack FP32 parameters into the FP8 record (one block), in the base-2 units
// of mlp_act (the hidden weights are unchanged); sw2, sw3: the hidden
// layers' weight scales (largest |w| / kFp8Max, computed by the caller over
// each layer).
__device__ inline void mlp_pack8(const float* __restrict__ p, const float* __restrict__ norm,
                                 float sw2, float sw3, MlpPacked8* __restrict__ out, int tid,
                                 int nthr) {
    for (int k = tid; k < 4 * 64; k += nthr) {
        const int To = k / 64, l = k % 64, g = l / 16, o = 16 * To + l % 16;
        mlp_bf16x4 v;
        for (int j = 0; j < 4; ++j) {
            const float w = kMlpNegLog2e * p[kOffW1 + o * kMlpIn + j];
            const uint16_t hi = mlp_bf16(w), lo = mlp_bf16(w - mlp_from_bf16(hi));
            v[j] = (short)(g <= 1 ? hi : (g == 2 ? lo : 0));
        }
        out->w1[To][l] = v;
    }
    for (int k = tid; k < 2 * 8 * 64; k += nthr) {
        const int layer = k / (8 * 64), r = k % (8 * 64);
        const int t = r / 64, l = r % 64, To = t / 2, Ti = t % 2;
        const float* W = p + (layer == 0 ? kOffW2 : kOffW3);
        const float isw = 1.0f / (layer == 0 ? sw2 : sw3);
        float v[8];
        for (int j = 0; j < 8; ++j)
            v[j] = W[(16 * To + l % 16) * kMlpH + mlp_kslot(Ti, l / 16, j)] * isw;
        (layer == 0 ? out->w2 : out->w3)[t][l] = mlp_fp8_pack8(v);
    }
    for (int k = tid; k < kMlpH; k += nthr) {
        out->b1[k] = kMlpNegLog2e * p[kOffB1 + k];
        out->b2[k] = kMlpNegLog2e * p[kOffB2 + k];
        out->b3[k] = kMlpNegLog2e * p[kOffB3 + k];
        out->w4[k] = kMlpNegLn2 * p[kOffW4 + k];
    }
    if (tid == 0) {
        out->b4 = p[kOffB4];
        for (int k = 0; k < kMlpNorm; ++k) out->norm[k] = norm[k];
        out->sw2 = sw2;
        out->sw3 = sw3;
        for (int k = 0; k < 3; ++k) out->pad[k] = 0.0f;
    }
}

// Largest |x| over the wave (every lane gets it).
__device__ inline float mlp_wave_absmax(float x) {
    x = fabsf(x);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
    return x;
}

// One hidden layer on the FP8 matrix cores: h (this lane's 16 inputs: tile
// To's outputs 4 (l / 16) + i of the previous layer) quantized at the
// wave's scale, times the layer's e4m3 weights, rescaled, plus the bias.
__device__ inline void mlp_fp8_layer(const mlp_fp8x8 (*w)[64], float sw, const float* bias,
                                     const float (&h)[4][4], float (&out)[4][4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    float m = 0.0f;
#pragma unroll
    for (int To = 0; To < 4; ++To)
#pragma unroll
        for (int i = 0; i < 4; ++i) m = fmaxf(m, fabsf(h[To][i]));
    m = mlp_wave_absmax(m);
    const float sh = m > 0.0f ? m / kFp8Max : 1.0f;
    const float ish = 1.0f / sh;
    // a 32-wide input chunk Ti is the previous layer's tiles 2 Ti, 2 Ti + 1
    mlp_fp8x8 X[2];
#pragma unroll
    for (int Ti = 0; Ti < 2; ++Ti) {
        float v[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v[i] = h[2 * Ti][i] * ish;
            v[4 + i] = h[2 * Ti + 1][i] * ish;
        }
        X[Ti] = mlp_fp8_pack8(v);
    }
    const float sc = sw * sh;
#pragma unroll
    for (int To = 0; To < 4; ++To) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int Ti = 0; Ti < 2; ++Ti)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(w[To * 2 + Ti][lane], X[Ti], acc, 0,
                                                             0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) out[To][i] = fmaf(acc[i], sc, bias[16 * To + 4 * g + i]);
    }
}

// FP8 inference of 16 samples per wave (as mlp_infer16).
__device__ inline float mlp_infer16_fp8(const MlpPacked8* __restrict__ m, const float (&x)[4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    mlp_bf16x4 in = mlp_input_parts(m->norm, x, g);
    float h[4][4], z[4][4];
#pragma unroll
    for (int To = 0; To < 4; ++To) {
        mlp_f32x4 acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = m->b1[16 * To + 4 * g + i];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m->w1[To][lane], in, acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) h[To][i] = mlp_act(acc[i]);
    }
    mlp_fp8_layer(m->w2, m->sw2, m->b2, h, z);
#pragma unroll
    for (int To = 0; To < 4; ++To)
#pragma unroll
        for (int i = 0; i < 4; ++i) h[To][i] = mlp_act(z[To][i]);
    mlp_fp8_layer(m->w3, m->sw3, m->b3, h, z);
    float y = 0.0f;
#pragma unroll
    for (int To = 0; To < 4; ++To)
#pragma unroll
        for (int i = 0; i < 4; ++i) y = fmaf(m->w4[16 * To + 4 * g + i], mlp_act(z[To][i]), y);
    y += __shfl_xor(y, 16);
    y += __shfl_xor(y, 32);
    return fmaf(y + m->b4, m->norm[8], m->norm[9]);
}

}  // namespace xts
