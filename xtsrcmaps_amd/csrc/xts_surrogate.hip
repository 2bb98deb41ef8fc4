// xts_surrogate.hip -- the learned surrogate of the far-pixel cubature
// (no counterpart in the reference, whose pixel cubature is
// mm_cubature/convolve_pixel_psf_logsep.hpp:23-96 for every pixel).
//
//   sample_kernel     training samples from the FP64 cubature itself: random
//                     (source, pixel) pairs of an image, classified as the
//                     row-weight pass classifies them, features and target
//                     B1 - |A| from the rule's points in FP64
//   mlp_train_kernel  FP32 forward (and backward) of a minibatch, gradient
//                     of the mean squared error summed per block
//   adam_kernel       the Adam update; clears the gradient for the next step
//   pack_kernel       FP32 parameters -> the BF16 matrix-core record
//   infer_kernel      BF16 inference on v_mfma_f32_16x16x16_bf16, 16
//                     samples per wave
// The network and its features are described in xts_mlp.hpp.
#include <cmath>
#include <cstring>
#include <vector>

#include "xts_internal.hpp"
#include "xts_mlp.hpp"

namespace xts {

// ------------------------------------------------------------ random numbers
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
__device__ inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// ------------------------------------------------------------ training data
// One candidate per thread: a source uniform in FITS pixel coordinates over
// the image widened by a quarter on each side, a pixel uniform in the image.
// The pixel is a sample when the row-weight pass would list it for the
// network: its points' separation bounds (pixel_geom's, in FP64) span
// exactly one table node, inside [1e-4, 28) degrees.  Output row (8 floats):
// sigma_c, g_u, g_v, kappa, y = B1 - |A|, A, separation of the centre (deg),
// weight total; rows are written compactly through *count.
__device__ inline int sep_upper_index(double s) {
    if (s < 1e-4) return 2;
    const double j = 1.0 + kRecipStep * (kLn1e4 + log(s));
    const int i = 1 + (int)j;
    return i > 400 ? 400 : i;
}

__global__ void __launch_bounds__(256)
sample_kernel(Wcs w, int H, int W, const double* __restrict__ rule, int N, double rho_scale,
              uint64_t seed, long long n_cand, long long cap, float* __restrict__ out,
              unsigned long long* __restrict__ count) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cand) return;
    const uint64_t h0 = splitmix64(seed ^ splitmix64((uint64_t)c * 4 + 1));
    const uint64_t h1 = splitmix64(h0 + 1), h2 = splitmix64(h0 + 2), h3 = splitmix64(h0 + 3);
    const double p1 = -0.25 * H + 1.5 * H * u01(h0) + 0.5;
    const double p2 = -0.25 * W + 1.5 * W * u01(h1) + 0.5;
    const int ph = (int)(u01(h2) * H), pw = (int)(u01(h3) * W);
    double src[3];
    if (wcs_pix2dir(w, p1, p2, src) != 0) return;
    // corners LL, LR, UR, UL of pixel (ph, pw) as the row-weight pass orders
    // them: FITS (h + 0.5, w + 0.5), (h + 0.5, w + 1.5), (h + 1.5, w + 1.5),
    // (h + 1.5, w + 0.5)
    const double cp[4][2] = {{ph + 0.5, pw + 0.5}, {ph + 0.5, pw + 1.5}, {ph + 1.5, pw + 1.5},
                             {ph + 1.5, pw + 0.5}};
    double D[4][3];
    for (int k = 0; k < 4; ++k) {
        double d[3];
        if (wcs_pix2dir(w, cp[k][0], cp[k][1], d) != 0) return;
        for (int q = 0; q < 3; ++q) D[k][q] = d[q] - src[q];
    }
    double A3[3], B3[3], C3[3], E3[3];
    for (int q = 0; q < 3; ++q) {
        A3[q] = 0.125 * ((D[0][q] + D[1][q]) + (D[2][q] + D[3][q]));
        B3[q] = 0.125 * ((D[1][q] - D[0][q]) + (D[2][q] - D[3][q]));
        C3[q] = 0.125 * ((D[3][q] - D[0][q]) + (D[2][q] - D[1][q]));
        E3[q] = 0.125 * ((D[0][q] - D[1][q]) + (D[2][q] - D[3][q]));
    }
    const double x0 = sqrt(A3[0] * A3[0] + A3[1] * A3[1] + A3[2] * A3[2]);
    double hr = 0.0;
    for (int k = 0; k < 4; ++k) {
        double e2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double e = 0.5 * D[k][q] - A3[q];
            e2 += e * e;
        }
        hr = fmax(hr, sqrt(e2));
    }
    hr = hr * rho_scale * 1.0001 + 1e-5 * x0;
    const double slo = 2.0 * kR2D * asin(fmax(x0 - hr, 0.0));
    const double shi = 2.0 * kR2D * asin(fmin(x0 + hr, 1.0));
    if (!(slo >= 1e-4 && shi < 28.0)) return;
    const int ilo = sep_upper_index(slo), ihi = sep_upper_index(shi);
    if (ihi != ilo + 1) return;
    const int r0 = min(max(ilo - 1, 0), kND - 4);
    const double step = kSepStep;
    const double s0 = 1e-4 * exp((r0 - 1) * step), s1 = 1e-4 * exp(r0 * step);
    auto sigma = [&](double u, double v) {
        double e2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double x = A3[q] + u * B3[q] + v * (C3[q] + u * E3[q]);
            e2 += x * x;
        }
        const double s = 2.0 * kR2D * asin(fmin(sqrt(e2), 1.0));
        return (s - s1) / (s1 - s0);
    };
    const double sc = sigma(0.0, 0.0);
    const double su = sigma(1.0, 0.0), sd = sigma(-1.0, 0.0);
    const double sv = sigma(0.0, 1.0), sw = sigma(0.0, -1.0);
    double A = 0.0, B1 = 0.0, wt = 0.0;
    for (int p = 0; p < N; ++p) {
        const double wq = 0.25 * rule[3 * p + 2];
        const double sg = sigma(rule[3 * p], rule[3 * p + 1]);
        A += wq * sg;
        B1 += wq * fabs(sg);
        wt += wq;
    }
    const unsigned long long k = atomicAdd(count, 1ull);
    if ((long long)k >= cap) return;
    float* o = out + 8 * k;
    o[0] = (float)sc;
    o[1] = (float)(0.5 * (su - sd));
    o[2] = (float)(0.5 * (sv - sw));
    o[3] = (float)(su + sd + sv + sw - 4.0 * sc);
    o[4] = (float)(B1 - fabs(A));
    o[5] = (float)A;
    o[6] = (float)(2.0 * kR2D * asin(fmin(x0, 1.0)));
    o[7] = (float)wt;
}

// ------------------------------------------------------------ FP32 training
// One block pass takes 64 samples.  LDS: the parameters (W2, W3 with a row
// stride of 65 so that lanes running over either index hit distinct banks),
// the inputs and, per layer, pre-activations z and activations h [64][64];
// the backward pass overwrites z with the layer's error delta.
constexpr int kTrainS = 64, kTrainThreads = 256;
constexpr int kWs = 65;   // row stride of W2, W3 in LDS
struct TrainLds {
    float W1[kMlpH * kMlpIn], b1[kMlpH], W2[kMlpH * kWs], b2[kMlpH], W3[kMlpH * kWs],
        b3[kMlpH], w4[kMlpH], b4[4];
    float x[kTrainS][kMlpIn];
    float yt[kTrainS], dy[kTrainS];
    float z1[kTrainS][kMlpH], h1[kTrainS][kMlpH], z2[kTrainS][kMlpH], h2[kTrainS][kMlpH],
        z3[kTrainS][kMlpH], h3[kTrainS][kMlpH];
};

__device__ inline float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ inline float silu32(float z) { return z * sigm(z); }
__device__ inline float dsilu32(float z) {
    const float g = sigm(z);
    return g * (1.0f + z * (1.0f - g));
}

// z[s][j] = b[j] + sum_k Wt(j, k) h[s][k] for the thread's 16 samples
template <int K, typename WF>
__device__ inline void dense_fwd(WF wf, const float* b, const float (*hin)[K], float (*z)[kMlpH],
                                 float (*hout)[kMlpH], int nvalid) {
    const int j = threadIdx.x % kMlpH, s0 = threadIdx.x / kMlpH;
    float acc[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) acc[m] = b[j];
    for (int k = 0; k < K; ++k) {
        const float wjk = wf(j, k);
#pragma unroll
        for (int m = 0; m < 16; ++m) acc[m] = fmaf(wjk, hin[s0 + 4 * m][k], acc[m]);
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int s = s0 + 4 * m;
        z[s][j] = acc[m];
        hout[s][j] = s < nvalid ? silu32(acc[m]) : 0.0f;
    }
}

// grad == nullptr: forward only, y_out [B] receives the predictions.
// Otherwise the samples' squared-error gradient (times inv_batch) is added
// to grad and their squared errors (times inv_batch) to loss.  With
// n_pool > 0 sample b of the batch is row splitmix64(seed, b) mod n_pool of
// X / Y, else row b.
__global__ void __launch_bounds__(kTrainThreads)
mlp_train_kernel(const float* __restrict__ P, const float* __restrict__ norm,
                 const float* __restrict__ X, int ldx, const float* __restrict__ Y, int ldy,
                 long long B, long long n_pool, uint64_t seed, float inv_batch,
                 float* __restrict__ grad, float* __restrict__ loss, float* __restrict__ y_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TrainLds& L = *reinterpret_cast<TrainLds*>(smem_raw);
    const int tid = threadIdx.x;
    for (int k = tid; k < kMlpH * kMlpIn; k += kTrainThreads) L.W1[k] = P[kOffW1 + k];
    for (int k = tid; k < kMlpH * kMlpH; k += kTrainThreads) {
        const int r = k / kMlpH, c = k % kMlpH;
        L.W2[r * kWs + c] = P[kOffW2 + k];
        L.W3[r * kWs + c] = P[kOffW3 + k];
    }
    for (int k = tid; k < kMlpH; k += kTrainThreads) {
        L.b1[k] = P[kOffB1 + k];
        L.b2[k] = P[kOffB2 + k];
        L.b3[k] = P[kOffB3 + k];
        L.w4[k] = P[kOffW4 + k];
    }
    if (tid == 0) L.b4[0] = P[kOffB4];
    const float ysc = norm[8], yoff = norm[9];
    // gradient sums of this block, per thread: j = tid % 64 and k = tid / 64
    // + 4 m for W2, W3; (j, i = tid / 64) for W1
    const int j = tid % kMlpH, q = tid / kMlpH;
    float gW2[16], gW3[16], gW1 = 0.f, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f,
          lsum = 0.f;
#pragma unroll
    for (int m = 0; m < 16; ++m) gW2[m] = gW3[m] = 0.f;
    const bool train = grad != nullptr;

    for (long long base = (long long)blockIdx.x * kTrainS; base < B;
         base += (long long)gridDim.x * kTrainS) {
        const int nvalid = (int)min((long long)kTrainS, B - base);
        __syncthreads();
        for (int k = tid; k < kTrainS; k += kTrainThreads) {
            long long row = base + k;
            if (k < nvalid && n_pool > 0)
                row = (long long)(splitmix64(seed ^ splitmix64((uint64_t)row)) % (uint64_t)n_pool);
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i)
                L.x[k][i] = k < nvalid ? (X[row * ldx + i] - norm[i]) * norm[4 + i] : 0.0f;
            L.yt[k] = (k < nvalid && Y) ? (Y[row * ldy] - yoff) / ysc : 0.0f;
        }
        __syncthreads();
        dense_fwd<kMlpIn>([&](int a, int b) { return L.W1[a * kMlpIn + b]; }, L.b1, L.x, L.z1,
                          L.h1, nvalid);
        __syncthreads();
        dense_fwd<kMlpH>([&](int a, int b) { return L.W2[a * kWs + b]; }, L.b2, L.h1, L.z2, L.h2,
                         nvalid);
        __syncthreads();
        dense_fwd<kMlpH>([&](int a, int b) { return L.W3[a * kWs + b]; }, L.b3, L.h2, L.z3, L.h3,
                         nvalid);
        __syncthreads();
        // output: four lanes per sample, 16 hidden units each
        {
            const int s = tid / 4, part = tid % 4;
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc = fmaf(L.w4[part * 16 + k], L.h3[s][part * 16 + k], acc);
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            if (part == 0) {
                const float yp = acc + L.b4[0];
                const float e = s < nvalid ? yp - L.yt[s] : 0.0f;
                L.dy[s] = 2.0f * e * inv_batch;
                lsum += e * e * inv_batch;
                if (!train && s < nvalid) y_out[base + s] = fmaf(yp, ysc, yoff);
            }
        }
        if (!train) continue;
        __syncthreads();
        // layer 4 and delta3 (into z3)
        {
            float a = 0.f, bsum = 0.f;
            for (int s = 0; s < kTrainS; ++s) {
                a = fmaf(L.dy[s], L.h3[s][j], a);
                bsum += L.dy[s];
            }
            if (q == 0) {
                gw4 += a;
                if (j == 0) gb4 += bsum;
            }
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int s = q + 4 * m;
                L.z3[s][j] = L.dy[s] * L.w4[j] * dsilu32(L.z3[s][j]);
            }
        }
        __syncthreads();
        // W3 gradient; delta2 = (W3^T delta3) silu'(z2) into z2
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int k = q + 4 * m;
            float a = 0.f;
            for (int s = 0; s < kTrainS; ++s) a = fmaf(L.z3[s][j], L.h2[s][k], a);
            gW3[m] += a;
        }
        if (q == 0) {
            float a = 0.f;
            for (int s = 0; s < kTrainS; ++s) a += L.z3[s][j];
            gb3 += a;
        }
        {
            const int k = j;   // lanes over the input index
            float acc[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) acc[m] = 0.f;
            for (int jj = 0; jj < kMlpH; ++jj) {
                const float wv = L.W3[jj * kWs + k];
#pragma unroll
                for (int m = 0; m < 16; ++m) acc[m] = fmaf(wv, L.z3[q + 4 * m][jj], acc[m]);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int s = q + 4 * m;
                L.z2[s][k] = acc[m] * dsilu32(L.z2[s][k]);
            }
        }
        __syncthreads();
        // W2 gradient; delta1 into z1
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int k = q + 4 * m;
            float a = 0.f;
            for (int s = 0; s < kTrainS; ++s) a = fmaf(L.z2[s][j], L.h1[s][k], a);
            gW2[m] += a;
        }
        if (q == 0) {
            float a = 0.f;
            for (int s = 0; s < kTrainS; ++s) a += L.z2[s][j];
            gb2 += a;
        }
        {
            const int k = j;
            float acc[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) acc[m] = 0.f;
            for (int jj = 0; jj < kMlpH; ++jj) {
                const float wv = L.W2[jj * kWs + k];
#pragma unroll
                for (int m = 0; m < 16; ++m) acc[m] = fmaf(wv, L.z2[q + 4 * m][jj], acc[m]);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int s = q + 4 * m;
                L.z1[s][k] = acc[m] * dsilu32(L.z1[s][k]);
            }
        }
        __syncthreads();
        // W1 gradient (j, i = q)
        {
            float a = 0.f, bsum = 0.f;
            for (int s = 0; s < kTrainS; ++s) {
                a = fmaf(L.z1[s][j], L.x[s][q], a);
                bsum += L.z1[s][j];
            }
            gW1 += a;
            if (q == 0) gb1 += bsum;
        }
    }
    if (train) {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int k = q + 4 * m;
            atomicAdd(&grad[kOffW2 + j * kMlpH + k], gW2[m]);
            atomicAdd(&grad[kOffW3 + j * kMlpH + k], gW3[m]);
        }
        atomicAdd(&grad[kOffW1 + j * kMlpIn + q], gW1);
        if (q == 0) {
            atomicAdd(&grad[kOffB1 + j], gb1);
            atomicAdd(&grad[kOffB2 + j], gb2);
            atomicAdd(&grad[kOffB3 + j], gb3);
            atomicAdd(&grad[kOffW4 + j], gw4);
            if (j == 0) atomicAdd(&grad[kOffB4], gb4);
        }
    }
    if (loss) {
        // the output lanes (tid % 4 == 0) hold partial sums
        lsum += __shfl_xor(lsum, 4);
        lsum += __shfl_xor(lsum, 8);
        lsum += __shfl_xor(lsum, 16);
        lsum += __shfl_xor(lsum, 32);
        if ((tid & 63) == 0) atomicAdd(loss, lsum);
    }
}

// Adam (Kingma & Ba 2015) with bias correction; the gradient is cleared for
// the next step's sums.
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = fmaf(b1, m[k], (1.0f - b1) * gk);
    const float vk = fmaf(b2, v[k], (1.0f - b2) * gk * gk);
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
    g[k] = 0.0f;
}

__global__ void fill_kernel(float* __restrict__ x, long long n, float val) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        x[k] = val;
}

// ------------------------------------------------------------ BF16 inference
template <int NP>
__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ P,
                                                   const float* __restrict__ norm,
                                                   MlpPackedT<NP>* __restrict__ out) {
    mlp_pack(P, norm, out, threadIdx.x, blockDim.x);
}

// The FP8 record: the hidden layers' weight scales from their largest
// |w|, then the packing (one block of 256 threads).
__global__ void __launch_bounds__(256) pack8_kernel(const float* __restrict__ P,
                                                    const float* __restrict__ norm,
                                                    MlpPacked8* __restrict__ out) {
    __shared__ float red[2][256];
    const int tid = threadIdx.x;
    float m2 = 0.0f, m3 = 0.0f;
    for (int k = tid; k < kMlpH * kMlpH; k += 256) {
        m2 = fmaxf(m2, fabsf(P[kOffW2 + k]));
        m3 = fmaxf(m3, fabsf(P[kOffW3 + k]));
    }
    red[0][tid] = m2;
    red[1][tid] = m3;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            red[0][tid] = fmaxf(red[0][tid], red[0][tid + o]);
            red[1][tid] = fmaxf(red[1][tid], red[1][tid + o]);
        }
        __syncthreads();
    }
    const float sw2 = red[0][0] > 0.0f ? red[0][0] / kFp8Max : 1.0f;
    const float sw3 = red[1][0] > 0.0f ? red[1][0] / kFp8Max : 1.0f;
    mlp_pack8(P, norm, sw2, sw3, out, tid, 256);
}

// Inference of B samples with the BF16 (MlpPacked) or FP8 (MlpPacked8)
// record, staged in LDS.
template <typename Rec>
__global__ void __launch_bounds__(256)
infer_kernel(const Rec* __restrict__ mp, const float* __restrict__ X, int ldx,
             long long B, float* __restrict__ y) {
    __shared__ Rec m;
    {
        const float4* src = reinterpret_cast<const float4*>(mp);
        float4* dst = reinterpret_cast<float4*>(&m);
        for (int k = threadIdx.x; k < (int)(sizeof(Rec) / 16); k += blockDim.x)
            dst[k] = src[k];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long waves = (long long)gridDim.x * (blockDim.x / 64);
    for (long long g = (long long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; g * 16 < B;
         g += waves) {
        const long long r = g * 16 + (lane & 15);
        const long long rr = r < B ? r : B - 1;
        float x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = X[rr * ldx + i];
        float out;
        if constexpr (std::is_same_v<Rec, MlpPacked8>) out = mlp_infer16_fp8(&m, x);
        else out = mlp_infer16(&m, x);
        if (lane < 16 && r < B) y[r] = out;
    }
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_surrogate_samples_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* rule, int64_t N,
                              uint64_t seed, int64_t n_candidates, int64_t capacity, float* out,
                              uint64_t* count) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(wcs && rule && out && count, "xts_surrogate_samples_dev: null pointer");
    XTS_CHECK(wcs->naxis1 > 0 && wcs->naxis2 > 0 && wcs->naxis1 <= 65536 &&
                  wcs->naxis2 <= 65536,
              "bad image size");
    XTS_CHECK(N > 0 && N <= 1024, "rule size %lld out of range (1..1024)", (long long)N);
    XTS_CHECK(n_candidates > 0 && n_candidates < (1LL << 40) && capacity > 0,
              "bad candidate count or capacity");
    // the rule is read from host memory here: its bilinear bound
    std::vector<double> r((size_t)N * 3);
    XTS_HIP(hipMemcpy(r.data(), rule, r.size() * 8, hipMemcpyDeviceToHost));
    double rs = 1.0;
    for (int64_t p = 0; p < N; ++p) {
        const double u = r[3 * p], v = r[3 * p + 1];
        rs = std::max(rs, 0.25 * (std::fabs((1 - u) * (1 - v)) + std::fabs((1 + u) * (1 - v)) +
                                  std::fabs((1 + u) * (1 + v)) + std::fabs((1 - u) * (1 + v))));
    }
    const Wcs w = wcs_setup(*wcs);
    sample_kernel<<<div_up(n_candidates, 256), 256, 0, ctx->stream>>>(
        w, (int)wcs->naxis1, (int)wcs->naxis2, rule, (int)N, rs, seed, (long long)n_candidates,
        (long long)capacity, out, (unsigned long long*)count);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

static int check_mlp_args(const float* params, const float* norm, const float* x, int64_t B) {
    XTS_CHECK(params && norm && x, "null device pointer");
    XTS_CHECK(B > 0 && B < (1LL << 40), "bad batch size %lld", (long long)B);
    return XTS_OK;
}

int xts_mlp_forward_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                        int64_t ldx, int64_t B, float* y) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_mlp_args(params, norm, x, B));
    XTS_CHECK(y && ldx >= kMlpIn, "null output or row stride below 4");
    XTS_TRY(allow_lds((const void*)mlp_train_kernel, sizeof(TrainLds)));
    const unsigned grid = std::min<unsigned>(div_up(B, kTrainS), 4096);
    mlp_train_kernel<<<grid, kTrainThreads, sizeof(TrainLds), ctx->stream>>>(
        params, norm, x, (int)ldx, nullptr, 1, B, 0, 0, 1.0f, nullptr, nullptr, y);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_mlp_grad_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                     int64_t ldx, const float* y_target, int64_t ldy, int64_t B, float* grad,
                     float* loss) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_mlp_args(params, norm, x, B));
    XTS_CHECK(y_target && grad && loss && ldx >= kMlpIn && ldy >= 1, "bad arguments");
    XTS_TRY(allow_lds((const void*)mlp_train_kernel, sizeof(TrainLds)));
    fill_kernel<<<div_up(kMlpParams, 256), 256, 0, ctx->stream>>>(grad, kMlpParams, 0.0f);
    fill_kernel<<<1, 64, 0, ctx->stream>>>(loss, 1, 0.0f);
    const unsigned grid = std::min<unsigned>(div_up(B, kTrainS), 4096);
    mlp_train_kernel<<<grid, kTrainThreads, sizeof(TrainLds), ctx->stream>>>(
        params, norm, x, (int)ldx, y_target, (int)ldy, B, 0, 0, 1.0f / (float)B, grad, loss,
        nullptr);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_mlp_adam_dev(xts_ctx* ctx, float* params, float* grad, float* m, float* v, int64_t step,
                     float lr, float beta1, float beta2, float eps) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(params && grad && m && v, "null device pointer");
    XTS_CHECK(step >= 1 && lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 &&
                  eps > 0,
              "bad Adam arguments");
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)step)));
    adam_kernel<<<div_up(kMlpParams, 256), 256, 0, ctx->stream>>>(params, grad, m, v, kMlpParams,
                                                                  lr, beta1, beta2, eps, c1, c2);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_surrogate_train_dev(xts_ctx* ctx, float* params, const float* norm, const float* x,
                            int64_t ldx, const float* y, int64_t ldy, int64_t n, int64_t steps,
                            int64_t batch, float lr, uint64_t seed, float* loss_hist) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_mlp_args(params, norm, x, n));
    XTS_CHECK(y && ldx >= kMlpIn && ldy >= 1 && steps > 0 && steps <= (1 << 24) && batch > 0 &&
                  batch <= (1 << 24) && lr > 0,
              "bad training arguments");
    XTS_TRY(allow_lds((const void*)mlp_train_kernel, sizeof(TrainLds)));
    void* st;
    XTS_TRY(slot_reserve(ctx, kSlotMlpState, (size_t)3 * kMlpParams * 4, &st));
    float* g = (float*)st;
    float* m = g + kMlpParams;
    float* v = m + kMlpParams;
    fill_kernel<<<div_up(3 * kMlpParams, 256), 256, 0, ctx->stream>>>(g, 3 * kMlpParams, 0.0f);
    if (loss_hist)
        fill_kernel<<<std::min<unsigned>(div_up(steps, 256), 1024), 256, 0, ctx->stream>>>(
            loss_hist, steps, 0.0f);
    const unsigned grid = std::min<unsigned>(div_up(batch, kTrainS), 4096);
    const double b1 = 0.9, b2 = 0.999;
    for (int64_t t = 1; t <= steps; ++t) {
        // cosine decay to a twentieth of lr
        const double frac = (double)(t - 1) / (double)steps;
        const float lrt = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
        mlp_train_kernel<<<grid, kTrainThreads, sizeof(TrainLds), ctx->stream>>>(
            params, norm, x, (int)ldx, y, (int)ldy, batch, n, splitmix64(seed + (uint64_t)t),
            1.0f / (float)batch, g, loss_hist ? loss_hist + (t - 1) : nullptr, nullptr);
        adam_kernel<<<div_up(kMlpParams, 256), 256, 0, ctx->stream>>>(
            params, g, m, v, kMlpParams, lrt, (float)b1, (float)b2, 1e-8f,
            (float)(1.0 / (1.0 - std::pow(b1, (double)t))),
            (float)(1.0 / (1.0 - std::pow(b2, (double)t))));
    }
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_mlp_infer_bf16_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                           int64_t ldx, int64_t B, float* y) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_mlp_args(params, norm, x, B));
    XTS_CHECK(y && ldx >= kMlpIn, "null output or row stride below 4");
    void* mp;
    XTS_TRY(slot_reserve(ctx, kSlotMlpPacked, sizeof(MlpPacked), &mp));
    pack_kernel<2><<<1, 256, 0, ctx->stream>>>(params, norm, (MlpPacked*)mp);
    const unsigned grid = std::min<unsigned>(div_up(B, 64), 4096);
    infer_kernel<MlpPacked><<<grid, 256, 0, ctx->stream>>>((const MlpPacked*)mp, x, (int)ldx, B,
                                                           y);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_mlp_infer_fp8_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                          int64_t ldx, int64_t B, float* y) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_mlp_args(params, norm, x, B));
    XTS_CHECK(y && ldx >= kMlpIn, "null output or row stride below 4");
    void* mp;
    XTS_TRY(slot_reserve(ctx, kSlotMlpPacked, sizeof(MlpPacked8), &mp));
    pack8_kernel<<<1, 256, 0, ctx->stream>>>(params, norm, (MlpPacked8*)mp);
    const unsigned grid = std::min<unsigned>(div_up(B, 64), 4096);
    infer_kernel<MlpPacked8><<<grid, 256, 0, ctx->stream>>>((const MlpPacked8*)mp, x, (int)ldx,
                                                            B, y);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_ctx_set_surrogate(xts_ctx* ctx, const float* params, const float* norm) {
    XTS_TRY(activate(ctx));
    if (!params) {
        ctx->surrogate_ready = false;
        return XTS_OK;
    }
    XTS_CHECK(norm, "null normalisation record");
    void *mp, *mph, *mp8;
    XTS_TRY(slot_reserve(ctx, kSlotSurrogate, sizeof(MlpPacked), &mp));
    XTS_TRY(slot_reserve(ctx, kSlotSurrogateH, sizeof(MlpPackedH), &mph));
    XTS_TRY(slot_reserve(ctx, kSlotSurrogate8, sizeof(MlpPacked8), &mp8));
    pack_kernel<2><<<1, 256, 0, ctx->stream>>>(params, norm, (MlpPacked*)mp);
    pack_kernel<1><<<1, 256, 0, ctx->stream>>>(params, norm, (MlpPackedH*)mph);
    pack8_kernel<<<1, 256, 0, ctx->stream>>>(params, norm, (MlpPacked8*)mp8);
    XTS_HIP(hipGetLastError());
    // packed before any later model map reads it, whatever stream that uses
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    ctx->surrogate = mp;
    ctx->surrogate_h = mph;
    ctx->surrogate8 = mp8;
    ctx->surrogate_ready = true;
    ++ctx->generation;
    return XTS_OK;
}

}  // extern "C"
