// xts_mfma.hpp -- FP64-in / FP64-out contraction on the BF16 matrix cores.
//
// Each FP64 operand is split into three BF16 terms x = hi + mid + lo (each
// term the BF16 rounding of the remaining residual, 8 significant bits
// apiece, ~24 bits together).  The product keeps the six terms down to
// 2^-24 relative (hi.hi, hi.mid, mid.hi, hi.lo, lo.hi, mid.mid) and
// accumulates in FP32 on v_mfma_f32_16x16x16_bf16: FP32-level accuracy at
// BF16 matrix-core throughput.
#pragma once

#include "xts_internal.hpp"

namespace xts {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline double bf16_to_double(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

__host__ __device__ inline void split3(double x, uint16_t* hi, uint16_t* mid, uint16_t* lo) {
    *hi = bf16_rne((float)x);
    const double r1 = x - bf16_to_double(*hi);
    *mid = bf16_rne((float)r1);
    const double r2 = r1 - bf16_to_double(*mid);
    *lo = bf16_rne((float)r2);
}

// A [M, K] row-major -> three [M, Kp] BF16 planes, Kp = K rounded up to 16,
// zero padded.  trans = 1 reads A as [K, M] (B operand stored column-major).
static __global__ void split3_kernel(const double* __restrict__ src, int rows, int K, int Kp, int trans,
                              uint16_t* __restrict__ out) {
    const size_t n = (size_t)rows * Kp;
    const size_t plane = n;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / Kp), c = (int)(k % Kp);
        double x = 0.0;
        if (c < K) x = trans ? src[(size_t)c * rows + r] : src[(size_t)r * K + c];
        uint16_t h, m, l;
        split3(x, &h, &m, &l);
        out[k] = h;
        out[plane + k] = m;
        out[2 * plane + k] = l;
    }
}

// One wave per 16x16 output tile, four waves per block.  Lane l holds A row
// (l % 16), k = 4 (l / 16) .. +3 of each 16-wide k step, B column (l % 16)
// likewise, and D rows 4 (l / 16) .. +3 of column (l % 16).
static __global__ void __launch_bounds__(256)
gemm_bf16x3_kernel(const uint16_t* __restrict__ A3, const uint16_t* __restrict__ B3,
                   double* __restrict__ C, int M, int N, int Kp, int tiles_n) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int tile = blockIdx.x * 4 + wave;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    if (tm * 16 >= M) return;
    const int row = tm * 16 + (lane % 16);
    const int col = tn * 16 + (lane % 16);
    const int kq = 4 * (lane / 16);
    const size_t pa = (size_t)M * Kp, pb = (size_t)N * Kp;
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const bool rv = row < M, cv = col < N;
    for (int k0 = 0; k0 < Kp; k0 += 16) {
        bf16x4_t ah = {0, 0, 0, 0}, am = ah, al = ah, bh = ah, bm = ah, bl = ah;
        if (rv) {
            const size_t o = (size_t)row * Kp + k0 + kq;
            ah = *reinterpret_cast<const bf16x4_t*>(A3 + o);
            am = *reinterpret_cast<const bf16x4_t*>(A3 + pa + o);
            al = *reinterpret_cast<const bf16x4_t*>(A3 + 2 * pa + o);
        }
        if (cv) {
            const size_t o = (size_t)col * Kp + k0 + kq;
            bh = *reinterpret_cast<const bf16x4_t*>(B3 + o);
            bm = *reinterpret_cast<const bf16x4_t*>(B3 + pb + o);
            bl = *reinterpret_cast<const bf16x4_t*>(B3 + 2 * pb + o);
        }
        // smallest terms first
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
    }
    const int ocol = tn * 16 + (lane % 16);
    if (ocol < N) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int orow = tm * 16 + 4 * (lane / 16) + r;
            if (orow < M) C[(size_t)orow * N + ocol] = (double)acc[r];
        }
    }
}

// [a; b]: stack two [S, C] row blocks into [2S, C]
static __global__ void pack_rows_kernel(int S, int C, const double* __restrict__ a,
                                        const double* __restrict__ b, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * S * C) return;
    out[k] = k < S * C ? a[k] : b[k - S * C];
}

// C [M, N] = A [M, K] . B [K, N], all FP64 device arrays.
inline int gemm_bf16x3(xts_ctx* ctx, const double* A, const double* B, double* C, int M, int K,
                       int N) {
    XTS_CHECK(M > 0 && K > 0 && N > 0, "gemm_bf16x3: bad sizes");
    const int Kp = (K + 15) / 16 * 16;
    void* w;
    const size_t na = (size_t)M * Kp * 3, nb = (size_t)N * Kp * 3;
    XTS_TRY(slot_reserve(ctx, kSlotWork1, (na + nb) * 2 + 256, &w));
    uint16_t* A3 = (uint16_t*)w;
    uint16_t* B3 = A3 + na;
    split3_kernel<<<std::min<unsigned>(div_up((int64_t)M * Kp, 256), 8192), 256, 0, ctx->stream>>>(
        A, M, K, Kp, 0, A3);
    split3_kernel<<<std::min<unsigned>(div_up((int64_t)N * Kp, 256), 8192), 256, 0, ctx->stream>>>(
        B, N, K, Kp, 1, B3);
    const int tiles_m = (M + 15) / 16, tiles_n = (N + 15) / 16;
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    gemm_bf16x3_kernel<<<div_up(tiles, 4), 256, 0, ctx->stream>>>(A3, B3, C, M, N, Kp, tiles_n);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int exposure_device(xts_ctx* ctx, const double* ecb, const double* wecb, const double* af,
                    const double* ab, const double* ltf, int64_t S, int64_t C, int64_t E,
                    double* out, int32_t precision);

}  // namespace xts
