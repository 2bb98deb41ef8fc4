// xts_exposure.hip -- (a) exposure on gfx950: effective-area bilinear
// sampling, HEALPix gather of the livetime cube, and the livetime x Aeff
// contraction with the livetime-efficiency scaling.
#include <algorithm>
#include <vector>

#include "xts_internal.hpp"
#include "xts_mfma.hpp"

namespace xts {

// math/bilerp.cxx:6-27: upper_bound bracket, (weight, complement, index);
// below the optional lower bound all weights are 0 and the index is 1.
// Returns false where the reference would read outside the grid.
__host__ __device__ inline bool lerp_pars(const double* g, int n, double v, bool zlb,
                                          double lb, double* wt, double* cw, int* idx) {
    if (zlb && v < lb) {
        *wt = 0.0; *cw = 0.0; *idx = 1;
        return true;
    }
    int up = 0;
    while (up < n && !(v < g[up])) ++up;
    if (up == 0 || up == n) return false;
    const double tt = (v - g[up - 1]) / (g[up] - g[up - 1]);
    *wt = tt; *cw = 1.0 - tt; *idx = up;
    return true;
}

// irf.aeff_value.cxx:11-55
__global__ void aeff_value_kernel(int C, int E, const double* __restrict__ costh,
                                  const double* __restrict__ logE, const double* __restrict__ pc,
                                  int Mc, const double* __restrict__ pe, int Me,
                                  const double* __restrict__ par, double minc,
                                  double* __restrict__ out, int* __restrict__ err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C * E) return;
    const int c = k / E, e = k % E;
    double cw, ccw, ew, ecw;
    int ci, ei;
    if (!lerp_pars(pc, Mc, costh[c], true, minc, &cw, &ccw, &ci) ||
        !lerp_pars(pe, Me, logE[e], false, 0.0, &ew, &ecw, &ei)) {
        atomicOr(err, 1);
        out[k] = 0.0;
        return;
    }
    const double xx = ccw * ecw * par[(ci - 1) * Me + ei - 1];
    const double xy = ccw * ew * par[(ci - 1) * Me + ei];
    const double yx = cw * ecw * par[ci * Me + ei - 1];
    const double yy = cw * ew * par[ci * Me + ei];
    out[k] = 1e4 * (xx + xy + yx + yy);
}

// HEALPix NESTED ang2pix (Gorski et al. 2005; healpix/healpix.cxx:182-229
// follows the same published algorithm).
__device__ inline uint64_t spread_bits(uint64_t v) {
    v &= 0xffffffffULL;
    v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
}

__device__ inline uint64_t ang2pix_nest(double theta, double phi, int64_t nside) {
    int order = 0;
    while ((1LL << order) < nside) ++order;
    const bool have_sth = (theta < 0.01) || (theta > kPi - 0.01);
    const double sth = have_sth ? sin(theta) : 0.0;
    const double z = cos(theta), za = fabs(z);
    double tt = phi * (2.0 / kPi);
    if (tt >= 0) { if (tt >= 4.0) tt = fmod(tt, 4.0); }
    else { tt = fmod(tt, 4.0) + 4.0; if (tt == 4.0) tt = 0.0; }
    uint64_t face, ix, iy;
    if (za <= 2.0 / 3.0) {
        const double t1 = nside * (0.5 + tt), t2 = nside * (z * 0.75);
        const uint64_t jp = (uint64_t)(t1 - t2), jm = (uint64_t)(t1 + t2);
        const uint64_t ifp = jp >> order, ifm = jm >> order;
        face = (ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : (ifm + 8));
        ix = jm & (nside - 1);
        iy = nside - (jp & (nside - 1)) - 1;
    } else {
        uint64_t ntt = (uint64_t)tt;
        if (ntt > 3) ntt = 3;
        const double tp = tt - ntt;
        const double tmp = ((za < 0.99) || !have_sth) ? nside * sqrt(3.0 * (1.0 - za))
                                                      : nside * sth / sqrt((1.0 + za) / 3.0);
        int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
        if (jp > nside - 1) jp = nside - 1;
        if (jm > nside - 1) jm = nside - 1;
        if (z >= 0) { face = ntt; ix = nside - jm - 1; iy = nside - jp - 1; }
        else { face = ntt + 8; ix = jp; iy = jm; }
    }
    return (face << (2 * order)) + (spread_bits(ix) | (spread_bits(iy) << 1));
}

// exposure.cxx:42-63: one thread per (source, cos-theta bin)
__global__ void src_cosbins_kernel(int S, int C, int64_t nside, const double* __restrict__ src,
                                   const double* __restrict__ cube, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S * C) return;
    const int s = k / C, c = k % C;
    const double theta = kHalfPi - src[2 * s + 1] * kD2R;
    const double phi = kD2R * src[2 * s];
    const uint64_t pix = ang2pix_nest(theta, phi, nside);
    out[k] = cube[pix * C + c];
}

// exposure.contract.cpp:13-42, FP64: one thread per output, K-loop in order
__global__ void contract_kernel(int M, int K, int N, const double* __restrict__ A,
                                const double* __restrict__ B, double* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (n >= N || m >= M) return;
    double acc = 0.0;
    for (int k = 0; k < K; ++k) acc = fma(A[(size_t)m * K + k], B[(size_t)k * N + n], acc);
    out[(size_t)m * N + n] = acc;
}

// exposure.cxx:96-169 fused: the four contractions, the front LTF rows and
// the sum ((lef + leb) + lwf) + lwb.
constexpr int kExpoThreads = 256, kExpoE = 64, kExpoG = kExpoThreads / kExpoE;
__global__ void __launch_bounds__(kExpoThreads)
exposure_kernel(int S, int C, int E, const double* __restrict__ ecb,
                const double* __restrict__ wecb, const double* __restrict__ af,
                const double* __restrict__ ab, const double* __restrict__ ltf,
                double* __restrict__ out) {
    // block: one source, kExpoE energies; kExpoG groups of threads take every
    // kExpoG-th cos theta bin, their partial sums are added in group order
    __shared__ double part[4][kExpoG][kExpoE];
    const int s = blockIdx.x, tx = threadIdx.x % kExpoE, g = threadIdx.x / kExpoE;
    const int e = blockIdx.y * kExpoE + tx;
    double ef = 0, wf = 0, eb = 0, wb = 0;
    if (e < E) {
        for (int c = g; c < C; c += kExpoG) {
            const double x = ecb[(size_t)s * C + c], y = wecb[(size_t)s * C + c];
            const double f = af[(size_t)c * E + e], b = ab[(size_t)c * E + e];
            ef = fma(x, f, ef);
            wf = fma(y, f, wf);
            eb = fma(x, b, eb);
            wb = fma(y, b, wb);
        }
    }
    part[0][g][tx] = ef;
    part[1][g][tx] = wf;
    part[2][g][tx] = eb;
    part[3][g][tx] = wb;
    __syncthreads();
    if (g == 0 && e < E) {
        for (int q = 1; q < kExpoG; ++q) {
            ef += part[0][q][tx];
            wf += part[1][q][tx];
            eb += part[2][q][tx];
            wb += part[3][q][tx];
        }
        const double l0 = ltf[e], l1 = ltf[E + e];
        out[(size_t)s * E + e] = ((ef * l0 + eb * l0) + wf * l1) + wb * l1;
    }
}

// BF16x3 variant: the sums over cos theta run on MFMA (see xts_mfma.hpp),
// then the same LTF scaling and summation order.
__global__ void exposure_combine_kernel(int S, int E, const double* __restrict__ P,
                                        const double* __restrict__ ltf, double* __restrict__ out) {
    // P = [ECB; WECB] (2S x C) . [Af | Ab] (C x 2E) -> 2S x 2E
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S * E) return;
    const int s = k / E, e = k % E;
    const size_t ld = 2 * (size_t)E;
    const double ef = P[(size_t)s * ld + e], eb = P[(size_t)s * ld + E + e];
    const double wf = P[(size_t)(S + s) * ld + e], wb = P[(size_t)(S + s) * ld + E + e];
    const double l0 = ltf[e], l1 = ltf[E + e];
    out[k] = ((ef * l0 + eb * l0) + wf * l1) + wb * l1;
}

__global__ void pack_cols_kernel(int C, int E, const double* __restrict__ a,
                                 const double* __restrict__ b, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C * 2 * E) return;
    const int c = k / (2 * E), j = k % (2 * E);
    out[k] = j < E ? a[(size_t)c * E + j] : b[(size_t)c * E + j - E];
}

int exposure_device(xts_ctx* ctx, const double* ecb, const double* wecb, const double* af,
                    const double* ab, const double* ltf, int64_t S, int64_t C, int64_t E,
                    double* out, int32_t precision) {
    XTS_CHECK(S > 0 && C > 0 && E > 0 && S * E < (1LL << 31), "bad sizes S=%lld C=%lld E=%lld",
              (long long)S, (long long)C, (long long)E);
    XTS_CHECK(precision == XTS_PREC_FP64 || precision == XTS_PREC_BF16X3, "bad precision %d",
              precision);
    if (precision == XTS_PREC_FP64) {
        XTS_CHECK(S <= 0x7fffffff && E <= 65535LL * kExpoE, "bad sizes");
        exposure_kernel<<<dim3((unsigned)S, (unsigned)div_up(E, kExpoE)), kExpoThreads, 0,
                          ctx->stream>>>((int)S, (int)C, (int)E, ecb, wecb, af, ab, ltf, out);
        XTS_HIP(hipGetLastError());
        return XTS_OK;
    }
    void* w;
    const size_t na = 2 * (size_t)S * C, nb = (size_t)C * 2 * E, np = 4 * (size_t)S * E;
    XTS_TRY(slot_reserve(ctx, kSlotWork0, (na + nb + np) * 8, &w));
    double* A = (double*)w;
    double* B = A + na;
    double* P = B + nb;
    pack_rows_kernel<<<div_up(na, 256), 256, 0, ctx->stream>>>((int)S, (int)C, ecb, wecb, A);
    pack_cols_kernel<<<div_up(nb, 256), 256, 0, ctx->stream>>>((int)C, (int)E, af, ab, B);
    XTS_TRY(gemm_bf16x3(ctx, A, B, P, (int)(2 * S), (int)C, (int)(2 * E)));
    exposure_combine_kernel<<<div_up(S * E, 256), 256, 0, ctx->stream>>>((int)S, (int)E, P, ltf,
                                                                          out);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_aeff_value(xts_ctx* ctx, const double* costh, int64_t C, const double* logE, int64_t E,
                   const xts_aeff_grid* g, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(costh && logE && g && g->params && g->cosths && g->logEs && out, "null pointer");
    XTS_CHECK(C > 0 && E > 0 && g->Mc > 1 && g->Me > 1, "bad sizes");
    void *dc, *de, *dpc, *dpe, *dp, *dout, *derr;
    XTS_TRY(to_device(ctx, kSlotIn0, costh, C * 8, &dc));
    XTS_TRY(to_device(ctx, kSlotIn1, logE, E * 8, &de));
    XTS_TRY(to_device(ctx, kSlotIn2, g->cosths, g->Mc * 8, &dpc));
    XTS_TRY(to_device(ctx, kSlotIn3, g->logEs, g->Me * 8, &dpe));
    XTS_TRY(to_device(ctx, kSlotIn4, g->params, g->Mc * g->Me * 8, &dp));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, C * E * 8, &dout));
    const int zero = 0;
    XTS_TRY(to_device(ctx, kSlotOut1, &zero, sizeof(int), &derr));
    aeff_value_kernel<<<div_up(C * E, 256), 256, 0, ctx->stream>>>(
        (int)C, (int)E, (double*)dc, (double*)de, (double*)dpc, (int)g->Mc, (double*)dpe,
        (int)g->Me, (double*)dp, g->min_costheta, (double*)dout, (int*)derr);
    XTS_HIP(hipGetLastError());
    int err = 0;
    XTS_HIP(hipMemcpyAsync(out, dout, C * E * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    if (err) return fail(XTS_ERR_RANGE, "xts_aeff_value: sample outside the IRF grid");
    return XTS_OK;
}

int xts_aeff_value_dev(xts_ctx* ctx, const double* costh, int64_t C, const double* logE,
                       int64_t E, const xts_aeff_grid* g, double* out, int32_t* range_error) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(costh && logE && g && g->params && g->cosths && g->logEs && out, "null pointer");
    XTS_CHECK(C > 0 && E > 0 && g->Mc > 1 && g->Me > 1 && C * E < (1LL << 31), "bad sizes");
    int* derr = (int*)range_error;
    if (!derr) {
        void* d;
        XTS_TRY(slot_reserve(ctx, kSlotOut1, sizeof(int), &d));
        derr = (int*)d;
    }
    aeff_value_kernel<<<div_up(C * E, 256), 256, 0, ctx->stream>>>(
        (int)C, (int)E, costh, logE, g->cosths, (int)g->Mc, g->logEs, (int)g->Me, g->params,
        g->min_costheta, out, derr);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_src_cosbins_dev(xts_ctx* ctx, const double* src_sph, int64_t S, const double* cube,
                        int64_t nside, int64_t C, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(src_sph && cube && out && S > 0 && C > 0 && S * C < (1LL << 31), "bad arguments");
    XTS_CHECK(nside > 0 && (nside & (nside - 1)) == 0 && nside <= (1LL << 29),
              "nside %lld is not a power of two", (long long)nside);
    src_cosbins_kernel<<<div_up(S * C, 256), 256, 0, ctx->stream>>>((int)S, (int)C, nside,
                                                                      src_sph, cube, out);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int xts_src_cosbins(xts_ctx* ctx, const double* src_sph, int64_t S, const double* cube,
                    int64_t nside, int64_t C, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(src_sph && cube && out && S > 0 && C > 0 && nside > 0, "bad arguments");
    void *ds, *dc, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, src_sph, S * 16, &ds));
    XTS_TRY(to_device(ctx, kSlotIn1, cube, 12 * nside * nside * C * 8, &dc));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, S * C * 8, &dout));
    XTS_TRY(xts_src_cosbins_dev(ctx, (double*)ds, S, (double*)dc, nside, C, (double*)dout));
    XTS_HIP(hipMemcpyAsync(out, dout, S * C * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_contract(xts_ctx* ctx, const double* A, const double* B, int64_t M, int64_t K, int64_t N,
                 double* out, int32_t precision) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(A && B && out && M > 0 && K > 0 && N > 0 && M < 65536, "bad arguments");
    XTS_CHECK(precision == XTS_PREC_FP64 || precision == XTS_PREC_BF16X3, "bad precision");
    void *da, *db, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, A, M * K * 8, &da));
    XTS_TRY(to_device(ctx, kSlotIn1, B, K * N * 8, &db));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, M * N * 8, &dout));
    if (precision == XTS_PREC_FP64) {
        contract_kernel<<<dim3(div_up(N, 256), (unsigned)M), 256, 0, ctx->stream>>>(
            (int)M, (int)K, (int)N, (double*)da, (double*)db, (double*)dout);
        XTS_HIP(hipGetLastError());
    } else {
        XTS_TRY(gemm_bf16x3(ctx, (double*)da, (double*)db, (double*)dout, (int)M, (int)K, (int)N));
    }
    XTS_HIP(hipMemcpyAsync(out, dout, M * N * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_exposure_dev(xts_ctx* ctx, const double* ecb, const double* wecb, const double* af,
                     const double* ab, const double* ltf, int64_t S, int64_t C, int64_t E,
                     double* out, int32_t precision) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(ecb && wecb && af && ab && ltf && out, "null device pointer");
    return exposure_device(ctx, ecb, wecb, af, ab, ltf, S, C, E, out, precision);
}

int xts_exposure(xts_ctx* ctx, const double* ecb, const double* wecb, const double* af,
                 const double* ab, const double* ltf, int64_t S, int64_t C, int64_t E,
                 double* out, int32_t precision) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(ecb && wecb && af && ab && ltf && out, "null host pointer");
    XTS_CHECK(S > 0 && C > 0 && E > 0, "bad sizes");
    void *d0, *d1, *d2, *d3, *d4, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, ecb, S * C * 8, &d0));
    XTS_TRY(to_device(ctx, kSlotIn1, wecb, S * C * 8, &d1));
    XTS_TRY(to_device(ctx, kSlotIn2, af, C * E * 8, &d2));
    XTS_TRY(to_device(ctx, kSlotIn3, ab, C * E * 8, &d3));
    XTS_TRY(to_device(ctx, kSlotIn4, ltf, 2 * E * 8, &d4));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, S * E * 8, &dout));
    XTS_TRY(exposure_device(ctx, (double*)d0, (double*)d1, (double*)d2, (double*)d3, (double*)d4,
                            S, C, E, (double*)dout, precision));
    XTS_HIP(hipMemcpyAsync(out, dout, S * E * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

}  // extern "C"
