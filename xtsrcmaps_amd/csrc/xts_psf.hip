// xts_psf.hip -- (b) mean-PSF build on gfx950: King PSF normalisation, the
// King table on the 401 separations, its bilinear sampling to the
// observation's (cos theta, log E) points weighted by the effective area, the
// livetime contraction, the mean over exposure and the FP64 radial integral.
#include <algorithm>
#include <vector>

#include "xts_internal.hpp"
#include "xts_mfma.hpp"

namespace xts {

// irf/irf.hxx:22-45 (gamma == 1 -> 1.001)
__host__ __device__ inline double king_f(double u, double g) {
    if (g == 1.0) g = 1.001;
    return (1.0 - 1.0 / g) * pow(1.0 + u / g, -g);
}

__host__ __device__ inline double evaluate_king(double sep, double sf, const double* p) {
    const double score = p[2] * sf, stail = p[3] * sf;
    const double rc = sep / score, rt = sep / stail;
    return p[0] * king_f(rc * rc / 2.0, p[4]) + p[1] * p[0] * king_f(rt * rt / 2.0, p[5]);
}

// irf/irf.hxx:48-77
__host__ __device__ inline double base_integral(double radius, double sf, const double* p) {
    const double score = p[2] * sf, stail = p[3] * sf;
    const double sep = radius * kD2R;
    const double rc = sep / score, rt = sep / stail;
    const double uc = rc * rc / 2.0, ut = rt * rt / 2.0;
    return p[0] * (1.0 - pow(1.0 + uc / p[4], 1.0 - p[4])) * kTwoPi * score * score +
           p[1] * p[0] * (1.0 - pow(1.0 + ut / p[5], 1.0 - p[5])) * kTwoPi * stail * stail;
}

// irf/irf.normalize_rpsf.cpp:11-61, one thread per (cos theta, energy) node
__global__ void normalize_rpsf_kernel(int Mc, int Me, double* __restrict__ params,
                                      const double* __restrict__ logEs, double sp0, double sp1,
                                      double si, const double* __restrict__ gl /* [64 x, 64 w] */) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Mc * Me) return;
    const int e = k % Me;
    double p[6];
    for (int i = 0; i < 6; ++i) p[i] = params[(size_t)k * 6 + i];
    const double energy = pow(10.0, logEs[e]);
    const double tt = pow(energy * 1.e-2, si);
    const double sf = sqrt(sp0 * tt * tt + sp1);
    double norm;
    if (energy < 120.0) {
        const double hw = 45.0, mid = 45.0;
        double acc = 0.0;
        for (int i = 0; i < 64; ++i) {
            const double v = hw * gl[i] + mid;
            const double kv = evaluate_king(v * kD2R, sf, p);
            acc += gl[64 + i] * (kv * (sin(v * kD2R) * kTwoPi * kD2R));
        }
        norm = acc * hw;
    } else {
        norm = base_integral(90.0, sf, p);
    }
    params[(size_t)k * 6 + 0] = p[0] / norm;
    params[(size_t)k * 6 + 2] = p[2] * sf;
    params[(size_t)k * 6 + 3] = p[3] * sf;
    params[(size_t)k * 6 + 4] = p[4] == 1.0 ? 1.001 : p[4];
    params[(size_t)k * 6 + 5] = p[5] == 1.0 ? 1.001 : p[5];
}

// psf/psf.king.cxx:9-43: K[d, c, e] at the separations (radians)
// Kernels over the PSF sections (front, back) take one argument set per
// section; blockIdx.y selects it, so both sections share one launch.
struct KingSec {
    const double* params;
    double* out;
    int Mc, Me;
};
struct KingSecs { KingSec s[2]; };

__device__ inline void king_table(const KingSec& q, const double* __restrict__ sep, int k0,
                                  int stride) {
    const int n = kND * q.Mc * q.Me;
    for (int k = k0; k < n; k += stride) {
        const int d = k / (q.Mc * q.Me), ce = k % (q.Mc * q.Me);
        q.out[k] = evaluate_king(sep[d] * kD2R, 1.0, q.params + (size_t)ce * 6);
    }
}

__global__ void king_kernel(KingSecs ks, const double* __restrict__ sep) {
    king_table(ks.s[blockIdx.y], sep, blockIdx.x * blockDim.x + threadIdx.x,
               gridDim.x * blockDim.x);
}

__host__ __device__ inline bool lerp_pars_psf(const double* g, int n, double v, double* wt,
                                              double* cw, int* idx) {
    int up = 0;
    while (up < n && !(v < g[up])) ++up;
    if (up == 0 || up == n) return false;
    const double tt = (v - g[up - 1]) / (g[up] - g[up - 1]);
    *wt = tt; *cw = 1.0 - tt; *idx = up;
    return true;
}

// psf/psf.bilerp.cxx:6-52: the bilinear weights and base node of each
// observation point (c, e), shared by all 401 separations.
struct BilerpPoint {
    double w00, w01, w10, w11;   // (1-tc)(1-te), (1-tc) te, tc (1-te), tc te
    int base;                    // (ci - 1) Me + ei - 1, or -1 outside the grid
    int pad;
};

struct BilerpSec {
    const double* pc;   // grid cos theta nodes [Mc]
    const double* pe;   // grid log E nodes [Me]
    BilerpPoint* pts;   // [C * E]
    int Mc, Me;
};
struct BilerpSecs { BilerpSec s[2]; };

__device__ inline void bilerp_point(const BilerpSec& g, int k, int C, int E,
                                    const double* __restrict__ costh,
                                    const double* __restrict__ logE, int* __restrict__ err) {
    if (k >= C * E) return;
    const int c = k / E, e = k % E;
    double cw, ccw, ew, ecw;
    int ci, ei;
    BilerpPoint q{};
    if (!lerp_pars_psf(g.pc, g.Mc, costh[c], &cw, &ccw, &ci) ||
        !lerp_pars_psf(g.pe, g.Me, logE[e], &ew, &ecw, &ei)) {
        atomicOr(err, 1);
        q.base = -1;
    } else {
        q.w00 = ccw * ecw;
        q.w01 = ccw * ew;
        q.w10 = cw * ecw;
        q.w11 = cw * ew;
        q.base = (ci - 1) * g.Me + ei - 1;
    }
    g.pts[k] = q;
}

__global__ void bilerp_points_kernel(int C, int E, const double* __restrict__ costh,
                                     const double* __restrict__ logE, BilerpSecs bs,
                                     int* __restrict__ err) {
    bilerp_point(bs.s[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x, C, E, costh, logE, err);
}

// Both of the above in one launch (they are independent): blocks
// [0, nking) of each section evaluate its King table, the rest its points.
__global__ void king_points_kernel(KingSecs ks, const double* __restrict__ sep, int nking, int C,
                                   int E, const double* __restrict__ costh,
                                   const double* __restrict__ logE, BilerpSecs bs,
                                   int* __restrict__ err) {
    if ((int)blockIdx.x < nking)
        king_table(ks.s[blockIdx.y], sep, blockIdx.x * blockDim.x + threadIdx.x,
                   nking * blockDim.x);
    else
        bilerp_point(bs.s[blockIdx.y], (blockIdx.x - nking) * blockDim.x + threadIdx.x, C, E,
                     costh, logE, err);
}

// times the effective area (psf.corrected_exposure_psf.cxx:37-44):
// PA[c, d, e] = bilerp(K)[c, d, e] x aeff[c, e] (aeff null: the bilinear
// samples alone).  One thread per (c, d, e); K is [401][Mc][Me].
struct AeffSec {
    const BilerpPoint* pts;
    const double* K;      // [401][Mc][Me]
    const double* aeff;   // [C][E] or null
    double* out;          // [C][401][E]
    int Mc, Me;
};
struct AeffSecs { AeffSec s[2]; };

__global__ void bilerp_aeff_kernel(int C, int E, AeffSecs as) {
    const AeffSec g = as.s[blockIdx.y];
    const size_t n = (size_t)C * kND * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % E);
        const int d = (int)((k / E) % kND);
        const int c = (int)(k / ((size_t)E * kND));
        const BilerpPoint q = g.pts[c * E + e];
        if (q.base < 0) {
            g.out[k] = 0.0;
            continue;
        }
        const double* Kd = g.K + (size_t)d * g.Mc * g.Me + q.base;
        const double v = q.w00 * Kd[0] + q.w01 * Kd[1] + q.w10 * Kd[g.Me] + q.w11 * Kd[g.Me + 1];
        g.out[k] = g.aeff ? v * g.aeff[(size_t)c * E + e] : v;
    }
}

// Both sections' PA summed, PA = PAf + PAb (the same livetime rows and, as
// the reference codes it, the front LTF for both: exposure.cxx:136-150,
// psf.psf_lookup_table_and_partial_integrals.cxx:42-47).
__global__ void bilerp_aeff_sum_kernel(int C, int E, AeffSecs as, double* __restrict__ out) {
    const size_t n = (size_t)C * kND * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % E);
        const int d = (int)((k / E) % kND);
        const int c = (int)(k / ((size_t)E * kND));
        double sum = 0.0;
#pragma unroll
        for (int sec = 0; sec < 2; ++sec) {
            const AeffSec& g = as.s[sec];
            const BilerpPoint q = g.pts[c * E + e];
            if (q.base < 0) continue;
            const double* Kd = g.K + (size_t)d * g.Mc * g.Me + q.base;
            const double v =
                q.w00 * Kd[0] + q.w01 * Kd[1] + q.w10 * Kd[g.Me] + q.w11 * Kd[g.Me + 1];
            sum += v * g.aeff[(size_t)c * E + e];
        }
        out[k] = sum;
    }
}

// psf.corrected_exposure_psf.cxx:46-117 and psf.mean_psf.cxx:20-30: for each
// source and column n = (d, e),
//   (ECB . PA) LTF0 + (WECB . PA) LTF1   (/ exposure when kMean),
// with PA = PAf + PAb for the mean PSF (front + back by linearity) or one
// section's PA.  FP64 FMA; a block covers kSrcTile sources x 256 columns,
// the livetime rows of its sources staged in LDS; a thread issues the loads
// of kCB cos-theta rows of PA before their FMAs.  Four-source tiles (twice
// the blocks of eight) ran 19.7 against 23.2 us for 100 sources.
constexpr int kSrcTile = 4;

template <bool kMean>
__global__ void __launch_bounds__(256)
mean_psf_kernel(int S, int C, int E, const double* __restrict__ ecb,
                const double* __restrict__ wecb, const double* __restrict__ pa,
                const double* __restrict__ ltf, const double* __restrict__ expo,
                double* __restrict__ out) {
    extern __shared__ double sh[];  // [2][kSrcTile][C]
    const int N = kND * E;
    const int s0 = blockIdx.y * kSrcTile;
    const int ns = min(kSrcTile, S - s0);
    for (int k = threadIdx.x; k < 2 * kSrcTile * C; k += blockDim.x) {
        const int which = k / (kSrcTile * C), r = (k / C) % kSrcTile, c = k % C;
        const double* src = which ? wecb : ecb;
        sh[k] = r < ns ? src[(size_t)(s0 + r) * C + c] : 0.0;
    }
    __syncthreads();
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int e = n % E;
    double ef[kSrcTile], wf[kSrcTile];
#pragma unroll
    for (int r = 0; r < kSrcTile; ++r) ef[r] = wf[r] = 0.0;
    constexpr int kCB = 20;
    for (int c0 = 0; c0 < C; c0 += kCB) {
        double fb[kCB];
#pragma unroll
        for (int j = 0; j < kCB; ++j) fb[j] = c0 + j < C ? pa[(size_t)(c0 + j) * N + n] : 0.0;
#pragma unroll
        for (int j = 0; j < kCB; ++j) {
            if (c0 + j >= C) break;
            const int c = c0 + j;
#pragma unroll
            for (int r = 0; r < kSrcTile; ++r) {
                ef[r] = fma(sh[r * C + c], fb[j], ef[r]);
                wf[r] = fma(sh[(kSrcTile + r) * C + c], fb[j], wf[r]);
            }
        }
    }
    const double l0 = ltf[e], l1 = ltf[E + e];
#pragma unroll
    for (int r = 0; r < kSrcTile; ++r) {
        if (r < ns) {
            const int s = s0 + r;
            const double v = ef[r] * l0 + wf[r] * l1;
            out[(size_t)s * N + n] = kMean ? v / expo[(size_t)s * E + e] : v;
        }
    }
}

// BF16x3-MFMA variant: P = [ECB; WECB] (2S x C) . PA (C x N) computed by
// gemm_bf16x3, then this combine.
__global__ void mean_psf_combine_kernel(int S, int E, const double* __restrict__ P,
                                        const double* __restrict__ ltf,
                                        const double* __restrict__ expo, double* __restrict__ out) {
    const int N = kND * E;
    const size_t n_all = (size_t)S * N;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_all;
         k += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(k / N), n = (int)(k % N), e = n % E;
        const double ef = P[(size_t)s * N + n], wf = P[(size_t)(S + s) * N + n];
        const double l0 = ltf[e], l1 = ltf[E + e];
        out[k] = (ef * l0 + wf * l1) / expo[(size_t)s * E + e];
    }
}

// psf.partial_total_integral.cxx:8-52 and psf.normalize.cxx:7-24 in FP64.
// Each segment d (nodes d-1, d) contributes the exact integral of the
// linear interpolant of y = 2 pi sin(theta) psf; the cumulative sum over d
// is normalised by its last value, and so is uPsf.  A block takes kIntE
// energies of one source; its threads split the 400 segments into kIntP
// contiguous parts (consecutive threads: consecutive energies, so every
// load is coalesced).  A thread loads its part's nodes and their theta and
// sin theta into registers at once and sums its segments; one thread per
// energy then forms the parts' exclusive prefixes and the total in LDS
// (summed in part order), and every thread writes its running sums and the
// normalised uPsf.  tab[d] = theta_d (rad), tab[401 + d] = sin theta_d.
// 16 parts of 25 segments: 14.9 us for 100 sources x 37 energies, against
// 17.4 us with 32 parts and 19.6 with 64.
constexpr int kIntE = 16, kIntP = 16;

__device__ inline double segment_integral(double t1, double t2, double s1, double s2, double u1,
                                          double u2) {
    const double y1 = kTwoPi * u1 * s1;
    const double y2 = kTwoPi * u2 * s2;
    const double m = (y2 - y1) / (t2 - t1);
    const double b = y1 - t1 * m;
    return 0.5 * m * (t2 * t2 - t1 * t1) + b * (t2 - t1);
}

constexpr int kIntSeg = (kND - 1 + kIntP - 1) / kIntP;   // segments of a part, at most

__global__ void __launch_bounds__(kIntE * kIntP)
integral_kernel(int S, int E, double* __restrict__ upsf, double* __restrict__ parint,
                double* __restrict__ total, const double* __restrict__ tab) {
    __shared__ double psum[kIntP][kIntE];
    __shared__ double pref[kIntP + 1][kIntE];   // exclusive prefixes; [kIntP]: total
    const int s = blockIdx.y;
    const int el = threadIdx.x % kIntE, part = threadIdx.x / kIntE;
    const int e = blockIdx.x * kIntE + el;
    const bool valid = e < E;
    const int ec = valid ? e : E - 1;
    double* u = upsf + (size_t)s * kND * E + ec;
    double* pi = parint + (size_t)s * kND * E + ec;
    constexpr int nseg = kND - 1;
    const int d_lo = 1 + part * nseg / kIntP, d_hi = 1 + (part + 1) * nseg / kIntP;
    const int n = d_hi - d_lo;
    // the part's nodes d_lo-1 .. d_hi-1 and their abscissae, loaded together
    double v[kIntSeg + 1], th[kIntSeg + 1], sn[kIntSeg + 1], seg[kIntSeg];
#pragma unroll
    for (int i = 0; i <= kIntSeg; ++i) {
        const int d = d_lo - 1 + min(i, n);
        v[i] = i <= n ? u[(size_t)d * E] : 0.0;
        th[i] = tab[d];
        sn[i] = tab[kND + d];
    }
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < kIntSeg; ++i) {
        seg[i] = i < n ? segment_integral(th[i], th[i + 1], sn[i], sn[i + 1], v[i], v[i + 1])
                       : 0.0;
        acc += seg[i];
    }
    psum[part][el] = acc;
    // after this barrier every thread holds its nodes: uPsf may be written
    __syncthreads();
    if (part == 0) {
        double run = 0.0;
        for (int q = 0; q < kIntP; ++q) {
            pref[q][el] = run;
            run += psum[q][el];
        }
        pref[kIntP][el] = run;
    }
    __syncthreads();
    if (!valid) return;
    const double tot = pref[kIntP][el];
    // running sums, normalised by the total (node 0 stays 0), and uPsf / total
    // (one reciprocal, then products)
    double run = pref[part][el];
    const double itot = 1.0 / tot;
    if (part == 0) {
        pi[0] = 0.0;
        u[0] = v[0] * itot;
        if (total) total[(size_t)s * E + e] = tot;
    }
#pragma unroll
    for (int i = 0; i < kIntSeg; ++i) {
        if (i < n) {
            run += seg[i];
            pi[(size_t)(d_lo + i) * E] = tot != 0.0 ? run * itot : run;
            u[(size_t)(d_lo + i) * E] = v[i + 1] * itot;
        }
    }
}

// Gauss-Legendre nodes and weights by Newton iteration on P_n.
static void legendre(int n, double* x, double* w) {
    for (int i = 1; i <= n; ++i) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= n; ++j) {
                const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) <= 1e-15) break;
        }
        x[i - 1] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

struct PsfDevGrid {
    const double* params;
    const double* cosths;
    const double* logEs;
    int Mc, Me;
};

// psf.corrected_exposure_psf.cxx:37-44: PA[c, d, e] = psf[c, d, e] aeff[c, e]
__global__ void psf_aeff_kernel(int C, int E, const double* __restrict__ psf,
                                const double* __restrict__ aeff, double* __restrict__ out) {
    const size_t n = (size_t)C * kND * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % E);
        const int c = (int)(k / ((size_t)E * kND));
        out[k] = psf[k] * aeff[(size_t)c * E + e];
    }
}

// psf.mean_psf.cxx:6-33: (front + back) / exposure[s, e]
__global__ void mean_kernel(int S, int E, const double* __restrict__ front,
                            const double* __restrict__ back, const double* __restrict__ expo,
                            double* __restrict__ out) {
    const size_t n = (size_t)S * kND * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % E);
        const size_t s = k / ((size_t)E * kND);
        out[k] = (front[k] + back[k]) / expo[s * E + e];
    }
}

// psf.normalize.cxx:7-24: uPsf[s, d, e] /= total[s, e]
__global__ void normalize_kernel(int S, int E, double* __restrict__ upsf,
                                 const double* __restrict__ total) {
    const size_t n = (size_t)S * kND * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % E);
        const size_t s = k / ((size_t)E * kND);
        upsf[k] /= total[s * E + e];
    }
}

static int psf_lookup_device(xts_ctx* ctx, const PsfDevGrid g[2], const double* costh, int C,
                             const double* logE, int E, const double* af, const double* ab,
                             const double* ecb, const double* wecb, const double* ltf,
                             const double* expo, int S, double* upsf, double* parint,
                             int32_t precision, int* err_dev) {
    XTS_CHECK(precision == XTS_PREC_FP64 || precision == XTS_PREC_BF16X3, "bad precision %d",
              precision);
    const int N = kND * E;
    const xts_const_tables* ct;
    XTS_TRY(const_tables(ctx, &ct));
    const double* dsep = ct->sep;
    // King tables and the aeff-weighted bilinear samples, front then back
    const size_t kf = (size_t)kND * g[0].Mc * g[0].Me, kb = (size_t)kND * g[1].Mc * g[1].Me;
    const size_t npa = (size_t)C * N;
    void* w;
    XTS_TRY(slot_reserve(ctx, kSlotWork0, (kf + kb + npa) * 8, &w));
    double* Kf = (double*)w;
    double* Kb = Kf + kf;
    double* PA = Kb + kb;   // front + back
    const KingSecs ks{{{g[0].params, Kf, g[0].Mc, g[0].Me}, {g[1].params, Kb, g[1].Mc, g[1].Me}}};
    void* bp;
    XTS_TRY(slot_reserve(ctx, kSlotBilerp, (size_t)2 * C * E * sizeof(BilerpPoint), &bp));
    BilerpPoint* bpf = (BilerpPoint*)bp;
    BilerpPoint* bpb = bpf + (size_t)C * E;
    const BilerpSecs bs{{{g[0].cosths, g[0].logEs, bpf, g[0].Mc, g[0].Me},
                         {g[1].cosths, g[1].logEs, bpb, g[1].Mc, g[1].Me}}};
    const int nking = (int)std::min<size_t>(div_up(std::max(kf, kb), 256), 4096);
    king_points_kernel<<<dim3(nking + div_up(C * E, 256), 2), 256, 0, ctx->stream>>>(
        ks, dsep, nking, C, E, costh, logE, bs, err_dev);
    const AeffSecs as{{{bpf, Kf, af, nullptr, g[0].Mc, g[0].Me},
                       {bpb, Kb, ab, nullptr, g[1].Mc, g[1].Me}}};
    bilerp_aeff_sum_kernel<<<std::min<unsigned>(div_up(npa, 256), 16384), 256, 0, ctx->stream>>>(
        C, E, as, PA);
    if (precision == XTS_PREC_FP64) {
        const size_t sm = 2 * kSrcTile * C * 8;
        mean_psf_kernel<true><<<dim3(div_up(N, 256), div_up(S, kSrcTile)), 256, sm, ctx->stream>>>(
            S, C, E, ecb, wecb, PA, ltf, expo, upsf);
    } else {
        void* w2;
        const size_t nA = 2 * (size_t)S * C, nP = 2 * (size_t)S * N;
        XTS_TRY(slot_reserve(ctx, kSlotWork2, (nA + nP) * 8, &w2));
        double* A = (double*)w2;
        double* P = A + nA;
        pack_rows_kernel<<<div_up(nA, 256), 256, 0, ctx->stream>>>(S, C, ecb, wecb, A);
        XTS_TRY(gemm_bf16x3(ctx, A, PA, P, 2 * S, C, N));
        mean_psf_combine_kernel<<<std::min<unsigned>(div_up((int64_t)S * N, 256), 16384), 256, 0,
                                  ctx->stream>>>(S, E, P, ltf, expo, upsf);
    }
    integral_kernel<<<dim3(div_up(E, kIntE), S), kIntE * kIntP, 0, ctx->stream>>>(
        S, E, upsf, parint, nullptr, ct->theta);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_normalize_rpsf(xts_ctx* ctx, double* params, const double* logEs, int64_t Mc, int64_t Me,
                       float scale0, float scale1, float scale_index) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(params && logEs && Mc > 0 && Me > 0, "bad arguments");
    double gl[128];
    legendre(64, gl, gl + 64);
    // IrfScale members are float: the squares are formed in float
    const float sp0 = scale0 * scale0, sp1 = scale1 * scale1;
    void *dp, *de, *dg;
    XTS_TRY(to_device(ctx, kSlotIn0, params, Mc * Me * 6 * 8, &dp));
    XTS_TRY(to_device(ctx, kSlotIn1, logEs, Me * 8, &de));
    XTS_TRY(to_device(ctx, kSlotIn2, gl, sizeof(gl), &dg));
    normalize_rpsf_kernel<<<div_up(Mc * Me, 64), 64, 0, ctx->stream>>>(
        (int)Mc, (int)Me, (double*)dp, (double*)de, (double)sp0, (double)sp1, (double)scale_index,
        (double*)dg);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(params, dp, Mc * Me * 6 * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_king(xts_ctx* ctx, const xts_psf_grid* g, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(g && g->params && out && g->Mc > 0 && g->Me > 0, "bad arguments");
    const xts_const_tables* ct;
    XTS_TRY(const_tables(ctx, &ct));
    void *dp, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, g->params, g->Mc * g->Me * 6 * 8, &dp));
    const size_t n = (size_t)kND * g->Mc * g->Me;
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 8, &dout));
    const KingSecs ks{{{(double*)dp, (double*)dout, (int)g->Mc, (int)g->Me}, {}}};
    king_kernel<<<dim3(std::min<unsigned>(div_up(n, 256), 4096), 1), 256, 0, ctx->stream>>>(
        ks, ct->sep);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

static int check_psf_args(const xts_psf_grid* f, const xts_psf_grid* b, int64_t C, int64_t E,
                          int64_t S) {
    XTS_CHECK(f && b && f->params && b->params && f->cosths && b->cosths && f->logEs && b->logEs,
              "null PSF grid");
    XTS_CHECK(f->Mc > 1 && f->Me > 1 && b->Mc > 1 && b->Me > 1, "bad PSF grid size");
    XTS_CHECK(C > 0 && E > 0 && S > 0 && (int64_t)kND * E < (1LL << 30) && S * E < (1LL << 31) &&
                  S * C < (1LL << 31),
              "bad sizes");
    return XTS_OK;
}

int xts_psf_lookup_table_and_partial_integrals_dev(
    xts_ctx* ctx, const xts_psf_grid* front, const xts_psf_grid* back, const double* costh,
    int64_t C, const double* logE, int64_t E, const double* af, const double* ab,
    const double* ecb, const double* wecb, const double* ltf, const double* expo, int64_t S,
    double* upsf, double* parint, int32_t precision) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_psf_args(front, back, C, E, S));
    XTS_CHECK(costh && logE && af && ab && ecb && wecb && ltf && expo && upsf && parint,
              "null device pointer");
    PsfDevGrid g[2] = {{front->params, front->cosths, front->logEs, (int)front->Mc, (int)front->Me},
                       {back->params, back->cosths, back->logEs, (int)back->Mc, (int)back->Me}};
    // range errors are reported by the host entry point; here they zero the
    // offending samples
    void* derr;
    XTS_TRY(slot_reserve(ctx, kSlotOut3, 256, &derr));
    return psf_lookup_device(ctx, g, costh, (int)C, logE, (int)E, af, ab, ecb, wecb, ltf, expo,
                             (int)S, upsf, parint, precision, (int*)derr);
}

int xts_psf_lookup_table_and_partial_integrals(
    xts_ctx* ctx, const xts_psf_grid* front, const xts_psf_grid* back, const double* costh,
    int64_t C, const double* logE, int64_t E, const double* af, const double* ab,
    const double* ecb, const double* wecb, const double* ltf, const double* expo, int64_t S,
    double* upsf, double* parint, int32_t precision) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_psf_args(front, back, C, E, S));
    XTS_CHECK(costh && logE && af && ab && ecb && wecb && ltf && expo && upsf && parint,
              "null host pointer");
    const xts_psf_grid* gs[2] = {front, back};
    PsfDevGrid g[2];
    for (int k = 0; k < 2; ++k) {
        void *dp, *dc, *de;
        XTS_TRY(to_device(ctx, kSlotIn0 + 3 * k, gs[k]->params, gs[k]->Mc * gs[k]->Me * 6 * 8, &dp));
        XTS_TRY(to_device(ctx, kSlotIn1 + 3 * k, gs[k]->cosths, gs[k]->Mc * 8, &dc));
        XTS_TRY(to_device(ctx, kSlotIn2 + 3 * k, gs[k]->logEs, gs[k]->Me * 8, &de));
        g[k] = {(double*)dp, (double*)dc, (double*)de, (int)gs[k]->Mc, (int)gs[k]->Me};
    }
    // small per-call inputs packed into one staging buffer
    const size_t nC = C, nE = E, nCE = C * E, nSC = S * C, nSE = S * E;
    std::vector<double> pack(nC + nE + 2 * nCE + 2 * nSC + 2 * nE + nSE);
    double* q = pack.data();
    std::copy(costh, costh + nC, q); q += nC;
    std::copy(logE, logE + nE, q); q += nE;
    std::copy(af, af + nCE, q); q += nCE;
    std::copy(ab, ab + nCE, q); q += nCE;
    std::copy(ecb, ecb + nSC, q); q += nSC;
    std::copy(wecb, wecb + nSC, q); q += nSC;
    std::copy(ltf, ltf + 2 * nE, q); q += 2 * nE;
    std::copy(expo, expo + nSE, q);
    void* dpk;
    XTS_TRY(to_device(ctx, kSlotIn6, pack.data(), pack.size() * 8, &dpk));
    const double* d = (const double*)dpk;
    const double *dcos = d, *dloge = dcos + nC, *daf = dloge + nE, *dab = daf + nCE,
                 *decb = dab + nCE, *dwecb = decb + nSC, *dltf = dwecb + nSC, *dexpo = dltf + 2 * nE;
    const size_t nout = (size_t)S * kND * E;
    void* dout;
    XTS_TRY(slot_reserve(ctx, kSlotOut0, nout * 8 * 2, &dout));
    double* dupsf = (double*)dout;
    double* dpar = dupsf + nout;
    const int zero = 0;
    void* derr;
    XTS_TRY(to_device(ctx, kSlotOut3, &zero, sizeof(int), &derr));
    XTS_TRY(psf_lookup_device(ctx, g, dcos, (int)C, dloge, (int)E, daf, dab, decb, dwecb, dltf,
                              dexpo, (int)S, dupsf, dpar, precision, (int*)derr));
    int err = 0;
    XTS_HIP(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(upsf, dupsf, nout * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(parint, dpar, nout * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    if (err) return fail(XTS_ERR_RANGE, "psf lookup: a sample lies outside the PSF IRF grid");
    return XTS_OK;
}

int xts_partial_total_integral(xts_ctx* ctx, double* upsf, int64_t S, int64_t E, double* parint,
                               double* total) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(upsf && parint && S > 0 && E > 0, "bad arguments");
    const size_t n = (size_t)S * kND * E;
    void *du, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, upsf, n * 8, &du));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 8 + S * E * 8, &dout));
    double* dpar = (double*)dout;
    double* dtot = dpar + n;
    const xts_const_tables* ct;
    XTS_TRY(const_tables(ctx, &ct));
    integral_kernel<<<dim3(div_up(E, kIntE), (unsigned)S), kIntE * kIntP, 0, ctx->stream>>>(
        (int)S, (int)E, (double*)du, dpar, dtot, ct->theta);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(upsf, du, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(parint, dpar, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (total) XTS_HIP(hipMemcpyAsync(total, dtot, S * E * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_bilerp(xts_ctx* ctx, const double* costh, int64_t C, const double* logE, int64_t E,
               const xts_psf_grid* grid, const double* kings, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(costh && logE && grid && grid->cosths && grid->logEs && kings && out, "null pointer");
    XTS_CHECK(C > 0 && E > 0 && grid->Mc > 1 && grid->Me > 1, "bad sizes");
    const size_t nk = (size_t)kND * grid->Mc * grid->Me, n = (size_t)C * kND * E;
    void *dc, *de, *dpc, *dpe, *dk, *dout, *derr;
    XTS_TRY(to_device(ctx, kSlotIn0, costh, C * 8, &dc));
    XTS_TRY(to_device(ctx, kSlotIn1, logE, E * 8, &de));
    XTS_TRY(to_device(ctx, kSlotIn2, grid->cosths, grid->Mc * 8, &dpc));
    XTS_TRY(to_device(ctx, kSlotIn3, grid->logEs, grid->Me * 8, &dpe));
    XTS_TRY(to_device(ctx, kSlotIn4, kings, nk * 8, &dk));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 8, &dout));
    const int zero = 0;
    XTS_TRY(to_device(ctx, kSlotOut3, &zero, sizeof(int), &derr));
    void* bp;
    XTS_TRY(slot_reserve(ctx, kSlotBilerp, (size_t)C * E * sizeof(BilerpPoint), &bp));
    const BilerpSecs bs{{{(double*)dpc, (double*)dpe, (BilerpPoint*)bp, (int)grid->Mc,
                          (int)grid->Me}, {}}};
    bilerp_points_kernel<<<dim3(div_up(C * E, 256), 1), 256, 0, ctx->stream>>>(
        (int)C, (int)E, (double*)dc, (double*)de, bs, (int*)derr);
    const AeffSecs as{{{(BilerpPoint*)bp, (double*)dk, nullptr, (double*)dout, (int)grid->Mc,
                        (int)grid->Me}, {}}};
    bilerp_aeff_kernel<<<dim3(std::min<unsigned>(div_up(n, 256), 16384), 1), 256, 0,
                         ctx->stream>>>((int)C, (int)E, as);
    XTS_HIP(hipGetLastError());
    int err = 0;
    XTS_HIP(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    if (err) return fail(XTS_ERR_RANGE, "bilerp: a sample lies outside the PSF IRF grid");
    return XTS_OK;
}

int xts_corrected_exposure_psf(xts_ctx* ctx, const double* obs_psf, const double* obs_aeff,
                               const double* ecb, const double* wecb, const double* front_ltf,
                               int64_t S, int64_t C, int64_t E, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(obs_psf && obs_aeff && ecb && wecb && front_ltf && out, "null pointer");
    XTS_CHECK(S > 0 && C > 0 && E > 0 && (int64_t)kND * E < (1LL << 30), "bad sizes");
    const size_t npa = (size_t)C * kND * E, nout = (size_t)S * kND * E;
    void *dpsf, *daeff, *decb, *dwecb, *dltf, *dpa, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, obs_psf, npa * 8, &dpsf));
    XTS_TRY(to_device(ctx, kSlotIn1, obs_aeff, C * E * 8, &daeff));
    XTS_TRY(to_device(ctx, kSlotIn2, ecb, S * C * 8, &decb));
    XTS_TRY(to_device(ctx, kSlotIn3, wecb, S * C * 8, &dwecb));
    XTS_TRY(to_device(ctx, kSlotIn4, front_ltf, 2 * E * 8, &dltf));
    XTS_TRY(slot_reserve(ctx, kSlotWork0, npa * 8, &dpa));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, nout * 8, &dout));
    psf_aeff_kernel<<<std::min<unsigned>(div_up(npa, 256), 16384), 256, 0, ctx->stream>>>(
        (int)C, (int)E, (double*)dpsf, (double*)daeff, (double*)dpa);
    const int N = kND * (int)E;
    const size_t sm = 2 * kSrcTile * C * 8;
    mean_psf_kernel<false><<<dim3(div_up(N, 256), div_up(S, kSrcTile)), 256, sm, ctx->stream>>>(
        (int)S, (int)C, (int)E, (double*)decb, (double*)dwecb, (double*)dpa, (double*)dltf,
        nullptr, (double*)dout);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, nout * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_mean_psf(xts_ctx* ctx, const double* front, const double* back, const double* exposure,
                 int64_t S, int64_t E, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(front && back && exposure && out && S > 0 && E > 0, "bad arguments");
    const size_t n = (size_t)S * kND * E;
    void *df, *db, *dx, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, front, n * 8, &df));
    XTS_TRY(to_device(ctx, kSlotIn1, back, n * 8, &db));
    XTS_TRY(to_device(ctx, kSlotIn2, exposure, S * E * 8, &dx));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 8, &dout));
    mean_kernel<<<std::min<unsigned>(div_up(n, 256), 16384), 256, 0, ctx->stream>>>(
        (int)S, (int)E, (double*)df, (double*)db, (double*)dx, (double*)dout);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_normalize(xts_ctx* ctx, double* upsf, const double* total, int64_t S, int64_t E) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(upsf && total && S > 0 && E > 0, "bad arguments");
    const size_t n = (size_t)S * kND * E;
    void *du, *dt;
    XTS_TRY(to_device(ctx, kSlotIn0, upsf, n * 8, &du));
    XTS_TRY(to_device(ctx, kSlotIn1, total, S * E * 8, &dt));
    normalize_kernel<<<std::min<unsigned>(div_up(n, 256), 16384), 256, 0, ctx->stream>>>(
        (int)S, (int)E, (double*)du, (double*)dt);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(upsf, du, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

}  // extern "C"
