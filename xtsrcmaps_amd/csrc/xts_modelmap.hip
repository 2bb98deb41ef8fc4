// xts_modelmap.hip -- (c) point-source model map on gfx950.
//
// Pipeline for one image (model_map.point_src_model_map_wcs.cxx:6-45):
//   1. geometry_kernel   FP64 pixel-corner and pixel-centre unit vectors of the
//                        WCS grid and the per-pixel solid angle (once per image)
//   2. radius_kernel     in-FOV flag and boundary radius per source
//   3. rowweight_kernel  per (source, pixel): the cubature reduced to four
//                        PSF-table row weights (x solid angle) and the
//                        in-radius row histogram of the map integral
//   4. near_rows_kernel  the same for the pixels whose rule points span more
//                        rows (near the source), up to 24 rows per pixel
//   5. hist_integral_kernel
//                        map integral, correction factor per (source, energy)
//   6. emit_*_kernel     every map value once: row weights . table rows
//                        x exposure x correction (the pooled near pixels as
//                        zeros); extra blocks evaluate the few pixels whose
//                        points touch more than 24 rows (wide_pixels)
//   7. near_emit_kernel  the pooled near pixels' values, over the zeros
#include <algorithm>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "xts_internal.hpp"
#include "xts_mlp.hpp"

namespace xts {

// ------------------------------------------------------------------ geometry
// corners[(i * (W+1) + j)*3]: direction of FITS pixel coordinate
// (i + 0.5, j + 0.5), i in [0, H], j in [0, W] -- the pixel corners of
// mm_cubature/points_weights.hpp:27-35.  centres: (i + 1, j + 1), the pixel
// centres used by math/solid_angle.hpp:50-75 and map_integral.cxx:29-31.
__global__ void geometry_kernel(Wcs w, int H, int W, double* __restrict__ corners,
                                double* __restrict__ centres) {
    const int n = (H + 1) * (W + 1);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = k / (W + 1), j = k % (W + 1);
        wcs_pix2dir(w, i + 0.5, j + 0.5, corners + 3 * k);
        wcs_pix2dir(w, i + 1.0, j + 1.0, centres + 3 * k);
    }
}

__device__ inline double vnorm3(const double* v) {
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// One flat triangle term of math/solid_angle.hpp:58-73.
__device__ inline double tri_omega(const double* A, const double* B, const double* D) {
    double u[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double v[3] = {A[0] - D[0], A[1] - D[1], A[2] - D[2]};
    const double nu = vnorm3(u), nv = vnorm3(v);
    for (int k = 0; k < 3; ++k) { u[k] /= nu; v[k] /= nv; }
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                         u[0] * v[1] - u[1] * v[0]};
    return dir_diff(A, B) * dir_diff(A, D) * vnorm3(c);
}

__global__ void solid_angle_kernel(int H, int W, const double* __restrict__ centres,
                                   double* __restrict__ omega) {
    const int n = H * W;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int h = k / W, w = k % W;
        const double* A = centres + 3 * (h * (W + 1) + w);
        const double* B = centres + 3 * (h * (W + 1) + w + 1);
        const double* C = centres + 3 * ((h + 1) * (W + 1) + w + 1);
        const double* D = centres + 3 * ((h + 1) * (W + 1) + w);
        omega[k] = 0.5 * (tri_omega(A, B, D) + tri_omega(C, B, D));
    }
}

// ------------------------------------------------------------------ radius
// model_map.psf_boundary_radius.cxx:9-62.  The 4 (N+1) boundary points at
// half-integer pixel coordinates are corner-grid nodes: (h+0.5, 0.5) is
// corner (h, 0), (h+0.5, W+0.5) is (h, W), (0.5, w+0.5) is (0, w) and
// (H+0.5, w+0.5) is (H, w).
struct SrcGeom {
    double dir[3];
    double radius_deg;
    double chord2_lo, chord2_hi;  // bracket of (2 sin(radius / 2))^2
    int in_fov;
    int pad;
};

// Block s < S: the boundary minimum of source s, taken on squared chords
// (the angle 2 asin(chord / 2) is increasing in the chord, so the minimum
// angle is the angle of the minimum chord); block S: the field-of-view flag
// of every source, one thread per source.
__device__ inline void radius_block(const Wcs& w, int H, int W, int S,
                                    const double* __restrict__ src_sph,
                                    const double* __restrict__ corners,
                                    SrcGeom* __restrict__ geo, int s) {
    if (s == S) {
        const double buf = 3.5;
        for (int t = threadIdx.x; t < S; t += blockDim.x) {
            double p1, p2;
            wcs_sph2pix(w, src_sph[2 * t], src_sph[2 * t + 1], &p1, &p2);
            geo[t].in_fov = p1 > buf && p1 < (H - buf) && p2 > buf && p2 < (W - buf);
        }
        return;
    }
    __shared__ double red[4];
    double src[3];
    sph2dir(src_sph[2 * s], src_sph[2 * s + 1], src);
    const int nb = 2 * (H + 1) + 2 * (W + 1);
    double mn = 16.0;   // above any squared chord of unit vectors
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        int ci, cj;
        if (k < H + 1) { ci = k; cj = 0; }
        else if (k < 2 * (H + 1)) { ci = k - (H + 1); cj = W; }
        else if (k < 2 * (H + 1) + W + 1) { ci = 0; cj = k - 2 * (H + 1); }
        else { ci = H; cj = k - 2 * (H + 1) - (W + 1); }
        const double* c = corners + 3 * (ci * (W + 1) + cj);
        const double x = src[0] - c[0], y = src[1] - c[1], z = src[2] - c[2];
        mn = fmin(mn, x * x + y * y + z * z);
    }
    for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)blockDim.x / 64; ++q) mn = fmin(mn, red[q]);
        // dir_diff of the nearest boundary point (sky_geom.hxx:305-310)
        const double rad = kR2D * (2.0 * asin(0.5 * sqrt(mn)));
        SrcGeom& g = geo[s];
        g.dir[0] = src[0]; g.dir[1] = src[1]; g.dir[2] = src[2];
        g.radius_deg = rad;
        const double c = 2.0 * sin(0.5 * kD2R * rad);
        g.chord2_lo = (c * (1.0 - 1e-12)) * (c * (1.0 - 1e-12));
        g.chord2_hi = (c * (1.0 + 1e-12)) * (c * (1.0 + 1e-12));
        g.pad = 0;
    }
}

__global__ void __launch_bounds__(256)
radius_kernel(Wcs w, int H, int W, int S, const double* __restrict__ src_sph,
              const double* __restrict__ corners, SrcGeom* __restrict__ geo) {
    radius_block(w, H, W, S, src_sph, corners, geo, blockIdx.x);
}

// ------------------------------------------------------------------ cubature
// The cubature of the reference (mm_cubature/convolve_pixel_psf_logsep.hpp:
// 20-80) interpolates the PSF table linearly in separation at every rule
// point of a pixel: map[e] = sum_p w_p ((1 - t_p) L[i_p - 1][e] + t_p L[i_p][e]).
// Regrouped by table row this is map[e] = sum_r W_r L[r][e]; every point of a
// pixel away from the source lands on one of four consecutive rows, so the
// pixel is described by four row weights and a first row.  The passes:
//   rowweight_kernel  point loop per (source, pixel): four row weights x
//                     solid angle, the in-radius row histogram of the map
//                     integral, and a list of the pixels whose points span
//                     more rows (near the source, or outside the table)
//   near_rows_kernel  up to 24 row weights per listed pixel (pooled), into
//                     the same histogram
//   emit_*_kernel     per (source, pixel, energy): the 4-row dot product x
//                     exposure x correction factor, stored once (pooled
//                     pixels as zeros); its extra blocks take the direct
//                     cubature of the few wide pixels (the source's own)
//   near_emit_kernel  the pooled pixels' values from their pool rows
constexpr int kTileW = 16, kTileH = 16;   // pixels per block pass (256 threads; a wave: 16x4)
constexpr int kThreads = kTileW * kTileH;
constexpr int kNearThreads = 256;
constexpr int kEmitThreads = 512;

// rule point: position (u, v) on [-1, 1]^2 and weight w/4
struct RuleUVW {
    double u, v, wq;
};

// rw_row of a listed pixel: kRowListed (set by the row-weight pass), turned
// into kRowWide by the near pass when the pixel is left to wide_pixels.
constexpr int kRowListed = -1, kRowWide = -2;

struct CubArgs {
    const double* corners;
    const double* centres;
    const double* omega;       // may be null (raw cubature)
    const SrcGeom* geo;
    const void* lut;           // [S][401][Ep] of T
    const RuleUVW* uvw;
    const float* lean;         // [3][N4] FP32 rule arrays u, v, w/4
    const float* inv_sep;      // [401] 2 R2D / sep[i] (sep in deg), FP32 table
    const double* inv_sep64;   // [401] FP64 table
    void* rw;                  // [S][H W] row weights (4 of T) x solid angle
    int* rw_row;               // [S][H W] first row; kRowListed, kRowWide for listed pixels
    double* hist;              // [S][401] in-radius row-weight sums, or null
    double* mapint;            // [S][E] in-radius map sums, or null (full model map only)
    int* near_count;           // [S]
    int* near_list;            // [S][H W] listed pixel indices
    int* near_off;             // [S][H W] pool offset per listed pixel (-1: wide)
    int* near_rr;              // [S][H W] rbase * 256 + R per listed pixel
    int* wide_list;            // [S][H W] list positions of the wide pixels
    int* wide_count;           // [S]
    void* near_pool;           // pooled row weights of listed pixels (T)
    int near_pool_cap;         // entries per source (kNearRows per list position)
    const double* expo;        // [S][E] or null
    const double* cf;          // [S][E] or null
    void* out;
    int S, H, W, E, Ep, N;
    int tiles_w, tiles_h, tiles_per_block;
    int rw_persist;            // row-weight pass on a 1-D grid of resident blocks
    int emit_gx;               // emit blocks per source (x)
    int emit_prio;             // wave priority of the emit's waves (s_setprio), 0: default
    int layout_sehw, out_f32;
    int tensor_nu, tensor_nv;  // rule is an nu x nv grid (p = i nv + j), else 0
    double rho_scale;          // max over points of sum_k |b_k|
    double wsum;               // sum over points of w/4
    double j0, jk;             // j = j0 + jk log2(asin(chord / 2))
    // learned surrogate of the one-node far pixels: 0 off, 1 the network,
    // 2 without it (|A| for the absolute moment)
    int sur;
    float4* sur_list;          // [S][tiles][256][2]: features; A, pixel, solid angle, 2 r0 + in-radius
    int* sur_count;            // [S][tiles] entries of each tile's 256 list slots
    const MlpPacked* sur_net;
    const MlpPacked8* sur_net8;
    const MlpPackedH* sur_neth;
    // summed model (xts_point_src_summed_model_dev): per-source flux and the
    // FP32 [E][H W] sum the batch's maps x flux are added to; else null
    const double* flux;
    float* sum;
    float* sum_fac;            // [S][Ep] flux x exposure x correction in FP32
};

__device__ inline float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994530942f; }
__device__ inline double fast_log(double x) { return log(x); }
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fast_sqrt(double x) { return sqrt(x); }
__device__ inline float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ inline double fast_log2(double x) { return log2(x); }

// upper node index of the reference's logsep lookup
// (convolve_pixel_psf_logsep.hpp:46-53): 2 below 1e-4 deg, clamped to 400.
template <typename T>
__device__ inline int sep_index_fast(T s) {
    if (s < T(1e-4)) return 2;
    const T j = T(1) + T(kRecipStep) * (T(kLn1e4) + fast_log(s));
    int i = 1 + (int)j;
    return i > 400 ? 400 : i;
}

template <typename T> struct Vec;
template <> struct Vec<float> { using type = float4; static constexpr int n = 4; };
template <> struct Vec<double> { using type = double2; static constexpr int n = 2; };
template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
constexpr int kNDPad = 404;   // 401 table entries, padded to keep 16-byte alignment

// Streaming (non-temporal) stores for the map: the map is written once and
// not re-read by the pass, and stores that bypass L2 allocation overlap far
// better with the LDS reads of the emit loop.
typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef double d4v_t __attribute__((ext_vector_type(4)));
__device__ inline void stream_store(float* p, float x) { __builtin_nontemporal_store(x, p); }
__device__ inline void stream_store(double* p, double x) { __builtin_nontemporal_store(x, p); }
__device__ inline void stream_store4(float* p, float a, float b, float c, float d) {
    __builtin_nontemporal_store(f4v_t{a, b, c, d}, reinterpret_cast<f4v_t*>(p));
}
__device__ inline void stream_store4(double* p, double a, double b, double c, double d) {
    __builtin_nontemporal_store(d4v_t{a, b, c, d}, reinterpret_cast<d4v_t*>(p));
}

template <typename V, typename T>
__device__ inline void vfma(V& acc, T w, const V& x);
template <>
__device__ inline void vfma<float4, float>(float4& a, float w, const float4& x) {
    a.x = fmaf(w, x.x, a.x); a.y = fmaf(w, x.y, a.y);
    a.z = fmaf(w, x.z, a.z); a.w = fmaf(w, x.w, a.w);
}
template <>
__device__ inline void vfma<double2, double>(double2& a, double w, const double2& x) {
    a.x = fma(w, x.x, a.x); a.y = fma(w, x.y, a.y);
}

// asin on [0, 1]: for x < 1/4 the odd Taylor series
// x + x^3 sum_k c_k x^(2k-2), c_k = (2k)! / (4^k (k!)^2 (2k+1)), through
// x^13 (the first omitted term is below 1e-10 relative); the library
// routine above 1/4 (separations beyond 29 degrees).
__device__ inline float asin_pos(float x) {
    if (x < 0.25f) {
        const float y = x * x;
        float p = 0.017352764423076924f;
        p = fmaf(p, y, 0.022372159090909092f);
        p = fmaf(p, y, 0.030381944444444444f);
        p = fmaf(p, y, 0.044642857142857144f);
        p = fmaf(p, y, 0.075f);
        p = fmaf(p, y, 0.16666666666666666f);
        return fmaf(x * y, p, x);
    }
    return asinf(x);
}
__device__ inline double asin_pos(double x) { return asin(x); }

// One far pixel's cubature point: half-chord x = |p - src| / 2 (the blend
// coefficients are pre-halved), a = asin(x), the lower table row
// jint = floor(j0 + jk log2 a) (the reference's upper index minus one) and
// t = redm1 (s / sep - 1) folded into wt = w redm1 (a isep[jint] - 1).
// Rows r0 .. r0+3 receive sums over the points of each relative row:
// S_r = sum w, T_r = sum w t.  NP: terms of the asin series (4 below 14
// degrees, 6 below 28 degrees).
template <typename T, int NP>
__device__ inline void far_point(T dx, T dy, T dz, T wq, T wqR, const T* __restrict__ isep_s,
                                 T j0, T jk, int r0, T& S0, T& T0, T& S1, T& T1, T& Tt) {
    const T y = dx * dx + dy * dy + dz * dz;
    const T x = fast_sqrt(y);
    T q;
    if (NP == 6) {
        q = T(0.017352764423076924);
        q = q * y + T(0.022372159090909092);
        q = q * y + T(0.030381944444444444);
    } else {
        q = T(0.030381944444444444);
    }
    q = q * y + T(0.044642857142857144);
    q = q * y + T(0.075);
    q = q * y + T(0.16666666666666666);
    const T a = sizeof(T) == 8 ? (T)asin((double)x) : x + (x * y) * q;
    const int jint = (int)(j0 + jk * fast_log2(a));
    const int rel = min(max(jint - r0, 0), 2);
    const T wt = wqR * (a * isep_s[jint] - T(1));
    const bool m0 = rel == 0, m1 = rel == 1;
    S0 += m0 ? wq : T(0);
    T0 += m0 ? wt : T(0);
    S1 += m1 ? wq : T(0);
    T1 += m1 ? wt : T(0);
    Tt += wt;
}

// Per-pixel set-up: corner offsets from the source (FP64, then T), half
// the bilinear blend p(u, v) / 2 = A + u B + v (C + u E) of the reference's
// point positions (points_weights.hpp:37-51), and bounds slo, shi (deg) on
// the separation of every point of the pixel.
template <typename T>
struct PixelGeom {
    T A[3], B[3], C[3], E[3];
    T slo, shi;
    T x0, hr;   // half-chord of the centre, bound on the points' offset from it
};

template <typename T>
__device__ inline PixelGeom<T> pixel_geom(const double* __restrict__ corners, int hc, int wc,
                                          int Wp1, const double* src, double rho_scale) {
    PixelGeom<T> g;
    const int c00 = hc * Wp1 + wc;
    const int cidx[4] = {c00, c00 + 1, c00 + Wp1 + 1, c00 + Wp1};  // LL LR UR UL
    T D[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double* c = corners + 3 * cidx[k];
        D[k][0] = (T)(c[0] - src[0]);
        D[k][1] = (T)(c[1] - src[1]);
        D[k][2] = (T)(c[2] - src[2]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g.A[c] = T(0.125) * ((D[0][c] + D[1][c]) + (D[2][c] + D[3][c]));
        g.B[c] = T(0.125) * ((D[1][c] - D[0][c]) + (D[2][c] - D[3][c]));
        g.C[c] = T(0.125) * ((D[3][c] - D[0][c]) + (D[2][c] - D[1][c]));
        g.E[c] = T(0.125) * ((D[0][c] - D[1][c]) + (D[2][c] - D[3][c]));
    }
    // half-chord bounds: |p - src| / 2 in [x0 - hr, x0 + hr]
    const T x0 = fast_sqrt(g.A[0] * g.A[0] + g.A[1] * g.A[1] + g.A[2] * g.A[2]);
    T hr = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const T ex = T(0.5) * D[k][0] - g.A[0], ey = T(0.5) * D[k][1] - g.A[1],
                ez = T(0.5) * D[k][2] - g.A[2];
        hr = fmax(hr, fast_sqrt(ex * ex + ey * ey + ez * ez));
    }
    hr = hr * (T)rho_scale * T(1.0001) + T(1e-5) * x0;
    g.x0 = x0;
    g.hr = hr;
    g.slo = T(2.0 * kR2D) * asin_pos(fmax(x0 - hr, T(0)));
    g.shi = T(2.0 * kR2D) * asin_pos(fmin(x0 + hr, T(1)));
    return g;
}

// Whether the pixel centre lies within the boundary radius
// (map_integral.cxx:29-36), decided on the chord with an exact check near
// the threshold.
// The exact test inside the bracket (rare), kept out of line so that it is
// branched to rather than evaluated for every pixel.
__device__ __attribute__((noinline)) bool in_radius_exact(double c2, double radius_deg) {
    return 2.0 * asin(0.5 * sqrt(c2)) <= kD2R * radius_deg;
}

__device__ inline bool in_radius(const SrcGeom& g, const double* __restrict__ centres, int idx) {
    const double* c = centres + 3 * idx;
    const double ex = g.dir[0] - c[0], ey = g.dir[1] - c[1], ez = g.dir[2] - c[2];
    const double c2 = ex * ex + ey * ey + ez * ez;
    // the bracket's 1e-12 margins dwarf the rounding of the squares
    if (c2 < g.chord2_lo) return true;
    if (c2 > g.chord2_hi) return false;
    return in_radius_exact(c2, g.radius_deg);
}

// Flat index of map value (s, pixel, e) in either layout.
__device__ inline size_t out_index(int layout_sehw, int s, int pix, int e, int E, size_t plane) {
    return layout_sehw ? ((size_t)s * E + e) * plane + pix : ((size_t)s * plane + pix) * E + e;
}

__device__ inline void store_value(const CubArgs& a, size_t k, double x) {
    if (a.out_f32) reinterpret_cast<float*>(a.out)[k] = (float)x;
    else reinterpret_cast<double*>(a.out)[k] = x;
}

// Scale factors of one energy chunk: x expo[s, e] then x correction[s, e],
// the order in which the reference scales the map
// (model_map.point_src_model_map_wcs.cxx:30-41); FP32 output takes one FP32
// multiply by the rounded product.
template <int EC>
struct ChunkScale {
    double ex[EC], cf[EC];
    float f32[EC];
    // threads 0 .. EC-1 load one energy each (both loads in flight together)
    __device__ void load(const CubArgs& a, int se0, int ne, int tid) {
        if (tid < EC) {
            const int k = se0 + min(tid, ne - 1);
            double x = a.expo[k], c = a.cf[k];
            if (tid >= ne) x = c = 1.0;
            ex[tid] = x;
            cf[tid] = c;
            // a summed model's values also carry the source's flux
            f32[tid] = a.flux ? (float)(x * c * a.flux[se0 / a.E]) : (float)(x * c);
        }
    }
    template <typename OT, typename T>
    __device__ OT apply(T x, int e) const {
        if constexpr (sizeof(OT) == 4) {
            return (float)x * f32[e];
        } else {
            double y = (double)x * ex[e];
            return y * cf[e];
        }
    }
};

template <typename T>
__device__ inline void stage_rule(const CubArgs& a, T* isep_s,
                                  typename Vec4<T>::type* rule_s, int N4, int tid, int nthr) {
    using V4 = typename Vec4<T>::type;
    for (int k = tid; k < kND; k += nthr)
        isep_s[k] = sizeof(T) == 4 ? (T)a.inv_sep[k] : (T)a.inv_sep64[k];
    for (int p = tid; p < N4; p += nthr) {
        V4 r = V4{T(0), T(0), T(0), T(0)};
        if (p < a.N) {
            const RuleUVW q = a.uvw[p];
            r = V4{(T)q.u, (T)q.v, (T)q.wq, (T)(q.wq * kRedm1)};
        }
        rule_s[p] = r;
    }
}

// PSF rows of source s for energies e0 .. e0+EC into [401][LS] (LS >= EC,
// zero-padded past the last energy).  All of a thread's loads are issued
// before the first LDS write so their latencies overlap.
template <typename T, int EC, int NTHR, int LS = EC>
__device__ inline void stage_lut(const CubArgs& a, int s, int e0, T* lut_s, int tid) {
    using V = typename Vec<T>::type;
    constexpr int VN = Vec<T>::n;
    constexpr int NV = EC / VN;
    constexpr int ITEMS = (kND * NV + NTHR - 1) / NTHR;
    const T* lut_g = reinterpret_cast<const T*>(a.lut) + (size_t)s * kND * a.Ep + e0;
    V buf[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int k = tid + i * NTHR;
        const int d = k / NV, v = k % NV;
        const bool ok = k < kND * NV && e0 + v * VN + VN <= a.Ep;
        // clamped, always in-bounds address; masked afterwards
        const int dc = ok ? d : 0, vc = ok ? v : 0;
        buf[i] = *reinterpret_cast<const V*>(lut_g + (size_t)dc * a.Ep + vc * VN);
        if (!ok) buf[i] = V{};
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int k = tid + i * NTHR;
        if (k >= kND * NV) break;
        const int d = k / NV, v = k % NV;
        if constexpr (LS == EC) {
            reinterpret_cast<V*>(lut_s)[d * NV + v] = buf[i];
        } else {
#pragma unroll
            for (int j = 0; j < VN; ++j)
                lut_s[d * LS + v * VN + j] = reinterpret_cast<const T*>(&buf[i])[j];
        }
    }
}

// ---- FP32 row and fraction of a point from the log of its separation (the
// near pass below 14 deg).
// With J(s) = ln(s / 1e-4 deg) / step the reference's lower row is
// 1 + floor(J) and t = (e^(step frac(J)) - 1) / (e^step - 1)
// (convolve_pixel_psf_logsep.hpp:46-54).  Per pixel r0 is its first row, so
// Q = 1 + J - r0 lies in [0, 3): rel = floor(Q), f = frac(Q).  With
// y = |p / 2|^2 (p: the un-normalised point, so s = 2 asin(sqrt y)) and
// z = y / y_c for the pixel centre's y_c,
//   Q = Q_c + K1 log2 z + g(y) / step,   K1 = ln 2 / (2 step),
//   g(y) = ln(asin(sqrt y) / sqrt y) = y (g1 + y (g2 + y (g3 + y g4))),
// Q_c = 1 + (ln(2 R2D) + ln 1e4 + ln(y_c) / 2) / step - r0 (per pixel, FP64).
// The relative log keeps Q accurate to ~1e-6 in FP32.  g is a least-squares
// fit on separations below 14.5 deg (cubic, max error 5e-11) or 28.5 deg
// (quartic, 1.2e-10; its derivatives feed the series of the row-weight
// pass); t(f) = f (t1 + f (t2 + f t3)) (max error 5e-8).
constexpr double kLnTwoR2D = 4.741374145600756;   // ln(2 R2D)
constexpr double kGCubic[3] = {0.16666666949401615, 0.06110791481077321, 0.03421911418748551};
constexpr double kGQuartic[4] = {0.1666666648047744, 0.06111208423547483, 0.03360670491188797,
                                 0.024068853992171683};
constexpr double kTPoly[3] = {0.9832291565474904, 0.016581961801669792, 0.00018883180009917163};

// z(u, v) = a(u) + v (b(u) + v c(u)), each of a, b, c quadratic in u, plus
// the Q offset and the cubic g's coefficients scaled to z: g_k y^k =
// g_k' z^k.
struct LeanPixel {
    float a0, a1, a2, b0, b1, b2, c0, c1, c2;
    float q0;
    float g[3];
};

// Natural log of a positive normal double: x = m 2^k with m in
// [sqrt(1/2), sqrt(2)), ln x = k ln 2 + 2 atanh(t), t = (m - 1) / (m + 1),
// |t| <= 0.1716; the atanh series to t^11 leaves < 3e-11 absolute.
__device__ inline double ln_pos(double x) {
    const long long bits = __double_as_longlong(x);
    int k = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((bits & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        ++k;
    }
    const double t = (m - 1.0) / (m + 1.0), t2 = t * t;
    double p = 1.0 / 11.0;
    p = fma(p, t2, 1.0 / 9.0);
    p = fma(p, t2, 1.0 / 7.0);
    p = fma(p, t2, 1.0 / 5.0);
    p = fma(p, t2, 1.0 / 3.0);
    p = fma(p, t2, 1.0);
    return fma((double)k, 0.69314718055994530942, 2.0 * t * p);
}

__device__ inline LeanPixel lean_pixel(const PixelGeom<float>& pg, int r0) {
    const float* A = pg.A; const float* B = pg.B; const float* C = pg.C; const float* E = pg.E;
    auto dot = [](const float* x, const float* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const float yc = dot(A, A);
    const float iy = 1.0f / yc;
    LeanPixel lp;
    lp.a0 = yc * iy;
    lp.a1 = 2.0f * dot(A, B) * iy;
    lp.a2 = dot(B, B) * iy;
    lp.b0 = 2.0f * dot(A, C) * iy;
    lp.b1 = 2.0f * (dot(A, E) + dot(B, C)) * iy;
    lp.b2 = 2.0f * dot(B, E) * iy;
    lp.c0 = dot(C, C) * iy;
    lp.c1 = 2.0f * dot(C, E) * iy;
    lp.c2 = dot(E, E) * iy;
    // log2 z = log2 y + log2 iy exactly as the kernel forms z = y iy
    double q0 = 1.0 + (kLnTwoR2D + kLn1e4) * kRecipStep -
                0.5 * kRecipStep * ln_pos((double)iy) - (double)r0;
    const double ycd = 1.0 / (double)iy;
    double ycp = ycd;
    for (int k = 0; k < 3; ++k) {
        lp.g[k] = (float)(kGCubic[k] * kRecipStep * ycp);
        ycp *= ycd;
    }
    lp.q0 = (float)q0;
    return lp;
}

// FP32 rule as arrays [N4] of u, v and w/4 in the constant address space:
// every lane of a wave reads the same point, so the reads are scalar loads
// (16 bytes per array for a group of four points) and the rule needs no
// copy in LDS.
using cfloat = const __attribute__((address_space(4))) float;
struct LeanRule {
    cfloat* p[3];   // u, v, wq
    const double* isep64;   // [401] 2 R2D / sep (global, or a copy in LDS)
};
constexpr int kLeanArrays = 3;

__device__ inline float4 lean_load4(cfloat* q) { return float4{q[0], q[1], q[2], q[3]}; }

__device__ inline float lane(const float4& x, int k) {
    return k == 0 ? x.x : (k == 1 ? x.y : (k == 2 ? x.z : x.w));
}

// ---- FP32 point loop without a logarithm, for pixels whose points lie
// close to the pixel centre relative to their separation.
// The row weights are linear in the separation s between rows (weight
// (s - s_{k-1}) / (s_k - s_{k-1}) to row k from below), so with rows r0,
// r0+1, r0+2, r0+3 at s0, s1 = s0 e^step, ... and
//   sigma = (s - s1) / (s1 - s0)
// they are linear combinations of 1, sigma, |sigma| and |sigma - G|,
// G = e^step = (s2 - s1) / (s1 - s0) (poly_rows).  sigma is a polynomial in
// x = y iy - 1 (iy ~ 1 / y_c, x quadratic in u and v with no constant term
// beyond the rounding of y_c iy).  With y_e = 1 / iy and s_e its separation,
//   s / s1 = (s_e / s1) sqrt(1 + x) exp(g(y_e (1 + x)) - g(y_e)),
// g as above; the Taylor series of the product to degree D times
// (s_e / s1) / (1 - e^-step), plus the constant (s_e / s1 - 1) /
// (1 - e^-step), is sigma(x); s_e / s1 - 1 is formed in FP64 from the
// arcsine series and the table of 2 R2D / sep.  A point costs the two FMAs
// of x, D of the polynomial and two (three with NM = 2) sums.
// Remainder: the binomial series of sqrt(1 + x) bounds it on |x| <= r by
// 1.1 K |binom(1/2, D + 1)| r^(D+1) / (1 - r), K = e^(2 step) /
// (1 - e^-step) = 32.25 (the exp factor's own terms add under 5%);
// kPolyReach[D - 2] is the r at which that bound reaches 1e-7 (in units of
// a row spacing, where the FP32 rounding of sigma is ~2e-7).
constexpr int kPolyMaxD = 6;
constexpr float kPolyReach[kPolyMaxD - 1] = {0.00355f, 0.0163f, 0.0397f, 0.0709f, 0.1065f};
constexpr double kRowG = 1.0 + 1.0 / kRedm1;           // e^step
constexpr double kInvD0 = kRedm1 + 1.0;                // 1 / (1 - e^-step)

struct PolyPixel {
    float x0, x1, x2;   // x(u, v) = x0 + u (x1 + u x2) + v (b(u) + v c(u))
    float b0, b1, b2, c0, c1, c2;
    float c[kPolyMaxD + 1];
};

// Bound r on |x| over the pixel's points: |p / 2| lies in [x0 - hr, x0 + hr],
// so z lies in [(1 - rho)^2, (1 + rho)^2] with rho = hr / x0.
__device__ inline float poly_reach(const PixelGeom<float>& pg) {
    const float rho = pg.hr * __builtin_amdgcn_rcpf(pg.x0) * 1.000001f;
    return rho * (2.0f + rho);
}

// asin(sqrt y) / sqrt y = sum a_n y^n, a_n = C(2n, n) / (4^n (2n + 1)); to
// y^8 the remainder is below 2e-13 on y <= sin^2(14.25 deg).
constexpr double kAsinSeries[9] = {1.0, 1.0 / 6, 3.0 / 40, 5.0 / 112, 35.0 / 1152, 63.0 / 2816,
                                   231.0 / 13312, 143.0 / 10240, 6435.0 / 557056};

template <int D>
__device__ inline PolyPixel poly_pixel(const double* isep64, const PixelGeom<float>& pg, int r0) {
    const float* A = pg.A; const float* B = pg.B; const float* C = pg.C; const float* E = pg.E;
    auto dot = [](const float* x, const float* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const float yc = dot(A, A);
    // any iy near 1 / yc serves: the series is expanded about y_e = 1 / iy
    const float iy = __builtin_amdgcn_rcpf(yc);
    PolyPixel pp;
    pp.x0 = fmaf(yc, iy, -1.0f);
    pp.x1 = 2.0f * dot(A, B) * iy;
    pp.x2 = dot(B, B) * iy;
    pp.b0 = 2.0f * dot(A, C) * iy;
    pp.b1 = 2.0f * (dot(A, E) + dot(B, C)) * iy;
    pp.b2 = 2.0f * dot(B, E) * iy;
    pp.c0 = dot(C, C) * iy;
    pp.c1 = 2.0f * dot(C, E) * iy;
    pp.c2 = dot(E, E) * iy;
    // s_e / s1 - 1 in FP64, s_e = 2 asin(sqrt y_e): y_e by one Newton step
    // on yc, sqrt y_e by one on the FP32 reciprocal square root
    const double yd = (double)yc;
    const double ye = yd * fma(-(double)iy, yd, 2.0);
    const double rs = (double)__builtin_amdgcn_rsqf(yc);
    double sq = ye * rs;
    sq = fma(fma(-sq, sq, ye), 0.5 * rs, sq);
    double G = kAsinSeries[8];
#pragma unroll
    for (int k = 7; k >= 0; --k) G = fma(G, ye, kAsinSeries[k]);
    // inv_sep64[k] = 2 R2D / sep[k]
    const double em = fma(sq * G, isep64[r0 + 1], -1.0);
    pp.c[0] = (float)(em * kInvD0);
    // exp(g(y_e (1 + x)) - g(y_e)) = 1 + e1 x + e2 x^2 + e3 x^3 + ..., times
    // K = (s_e / s1) / (1 - e^-step)
    const float K = (float)((1.0 + em) * kInvD0);
    const float yf = (float)ye;
    const float g1 = (float)kGQuartic[0], g2 = (float)kGQuartic[1], g3 = (float)kGQuartic[2],
                g4 = (float)kGQuartic[3];
    const float ga = yf * (g1 + yf * (2.0f * g2 + yf * (3.0f * g3 + yf * 4.0f * g4)));
    const float gb = yf * yf * (g2 + yf * (3.0f * g3 + yf * 6.0f * g4));
    const float gcb = yf * yf * yf * (g3 + yf * 4.0f * g4);
    const float ex[4] = {K, K * ga, K * (gb + 0.5f * ga * ga),
                         K * (gcb + ga * gb + ga * ga * ga * (1.0f / 6.0f))};
    // binomial series of sqrt(1 + x)
    constexpr float bs[kPolyMaxD + 1] = {1.0f, 0.5f, -0.125f, 0.0625f, -0.0390625f,
                                         0.02734375f, -0.0205078125f};
#pragma unroll
    for (int k = 1; k <= D; ++k) {
        float r = 0.0f;
#pragma unroll
        for (int j = 0; j <= 3 && j <= k; ++j) r = fmaf(bs[k - j], ex[j], r);
        pp.c[k] = r;
    }
    return pp;
}

// Row weights from A = sum wq sigma, B1 = sum wq |sigma|, B2 = sum wq
// |sigma - G| (wtot = sum wq): with P1 = sum wq max(sigma, 0) and P2 = sum
// wq max(sigma - G, 0), row r0 gets max(-sigma, 0), row r0+1 ... the hat
// of slopes 1 and 1 / G, row r0+2 max(sigma, 0) / G - max(sigma - G, 0)
// (1 / G + 1 / G^2), row r0+3 max(sigma - G, 0) / G^2.
template <int NM>
__device__ inline void poly_rows(float wtot, float A, float B1, float B2, float (&W)[4]) {
    constexpr float iG = (float)(1.0 / kRowG), iG2 = (float)(1.0 / (kRowG * kRowG));
    const float P1 = 0.5f * (B1 + A);
    W[0] = 0.5f * (B1 - A);
    if constexpr (NM < 2) {
        W[2] = P1 * iG;
        W[3] = 0.0f;
    } else {
        const float P2 = 0.5f * (B2 + fmaf(-(float)kRowG, wtot, A));
        W[2] = fmaf(P1, iG, -P2 * (iG + iG2));
        W[3] = P2 * iG2;
    }
    W[1] = wtot - W[0] - W[2] - W[3];
}

template <int D, int NM>
__device__ inline void poly_point(float x, float wq, const PolyPixel& pp, float& A, float& B1,
                                  float& B2) {
    float sg = pp.c[D];
#pragma unroll
    for (int k = D - 1; k >= 0; --k) sg = fmaf(sg, x, pp.c[k]);
    A = fmaf(wq, sg, A);
    B1 = fmaf(wq, fabsf(sg), B1);
    if constexpr (NM >= 2) B2 = fmaf(wq, fabsf(sg - (float)kRowG), B2);
}

template <int D, int NM>
__device__ inline void poly_tensor_loop(const CubArgs& a, const LeanRule& lr,
                                        const PixelGeom<float>& pg, int r0, float wtot,
                                        float (&W)[4]) {
    const PolyPixel pp = poly_pixel<D>(lr.isep64, pg, r0);
    const int nv = a.tensor_nv;
    float A = 0.0f, B1 = 0.0f, B2 = 0.0f;
    for (int i = 0; i < a.tensor_nu; ++i) {
        const float u = lr.p[0][i * nv];
        const float xu = fmaf(fmaf(pp.x2, u, pp.x1), u, pp.x0);
        const float bu = fmaf(fmaf(pp.b2, u, pp.b1), u, pp.b0);
        const float cu = fmaf(fmaf(pp.c2, u, pp.c1), u, pp.c0);
        for (int jj = 0; jj < nv; jj += 4) {
            const int p0 = i * nv + jj;
            const float4 v = lean_load4(lr.p[1] + p0), wq = lean_load4(lr.p[2] + p0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float vk = lane(v, k);
                poly_point<D, NM>(fmaf(fmaf(cu, vk, bu), vk, xu), lane(wq, k), pp, A, B1, B2);
            }
        }
    }
    poly_rows<NM>(wtot, A, B1, B2, W);
}

template <int D, int NM>
__device__ inline void poly_list_loop(const CubArgs& a, int N4, const LeanRule& lr,
                                      const PixelGeom<float>& pg, int r0, float wtot,
                                      float (&W)[4]) {
    const PolyPixel pp = poly_pixel<D>(lr.isep64, pg, r0);
    float A = 0.0f, B1 = 0.0f, B2 = 0.0f;
    for (int p0 = 0; p0 < N4; p0 += 4) {
        const float4 u = lean_load4(lr.p[0] + p0), v = lean_load4(lr.p[1] + p0),
                     wq = lean_load4(lr.p[2] + p0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float uk = lane(u, k), vk = lane(v, k);
            const float xu = fmaf(fmaf(pp.x2, uk, pp.x1), uk, pp.x0);
            const float bu = fmaf(fmaf(pp.b2, uk, pp.b1), uk, pp.b0);
            const float cu = fmaf(fmaf(pp.c2, uk, pp.c1), uk, pp.c0);
            poly_point<D, NM>(fmaf(fmaf(cu, vk, bu), vk, xu), lane(wq, k), pp, A, B1, B2);
        }
    }
    poly_rows<NM>(wtot, A, B1, B2, W);
}

template <int D, int NM>
__device__ inline void poly_loop(const CubArgs& a, const LeanRule& lr, const PixelGeom<float>& pg,
                                 int r0, int N4, float wtot, float (&W)[4]) {
    if (a.tensor_nu > 0) poly_tensor_loop<D, NM>(a, lr, pg, r0, wtot, W);
    else poly_list_loop<D, NM>(a, N4, lr, pg, r0, wtot, W);
}

// Runs the polynomial loop of the least degree whose reach covers every
// lane's rx; false when none does.
template <int NM>
__device__ inline bool poly_dispatch(const CubArgs& a, const LeanRule& lr,
                                     const PixelGeom<float>& pg, int r0, int N4, float wtot,
                                     float rx, float (&W)[4]) {
    if (__all(rx <= kPolyReach[0])) poly_loop<2, NM>(a, lr, pg, r0, N4, wtot, W);
    else if (__all(rx <= kPolyReach[1])) poly_loop<3, NM>(a, lr, pg, r0, N4, wtot, W);
    else if (__all(rx <= kPolyReach[2])) poly_loop<4, NM>(a, lr, pg, r0, N4, wtot, W);
    else if (__all(rx <= kPolyReach[3])) poly_loop<5, NM>(a, lr, pg, r0, N4, wtot, W);
    else if (__all(rx <= kPolyReach[4])) poly_loop<6, NM>(a, lr, pg, r0, N4, wtot, W);
    else return false;
    return true;
}

// ---- the surrogate's inputs for a one-node pixel (xts_mlp.hpp): sigma of
// the degree-D series at the centre and the four edge midpoints of the
// pixel's (u, v) square, and A from the 2x2 Gauss-Legendre rule of sigma
// (times the rule's weight total) in place of the point loop.
template <int D>
__device__ inline void sur_pixel(const double* isep64, const PixelGeom<float>& pg, int r0,
                                 float wsum, float (&f)[4], float& A) {
    const PolyPixel pp = poly_pixel<D>(isep64, pg, r0);
    auto sig = [&](float u, float v) {
        const float xu = fmaf(fmaf(pp.x2, u, pp.x1), u, pp.x0);
        const float bu = fmaf(fmaf(pp.b2, u, pp.b1), u, pp.b0);
        const float cu = fmaf(fmaf(pp.c2, u, pp.c1), u, pp.c0);
        const float x = fmaf(fmaf(cu, v, bu), v, xu);
        float sg = pp.c[D];
#pragma unroll
        for (int k = D - 1; k >= 0; --k) sg = fmaf(sg, x, pp.c[k]);
        return sg;
    };
    const float sc = sig(0.f, 0.f);
    const float su = sig(1.f, 0.f), sd = sig(-1.f, 0.f), sv = sig(0.f, 1.f), sw = sig(0.f, -1.f);
    f[0] = sc;
    f[1] = 0.5f * (su - sd);
    f[2] = 0.5f * (sv - sw);
    f[3] = (su + sd) + (sv + sw) - 4.0f * sc;
    constexpr float g = 0.57735026918962576f;   // 1 / sqrt(3)
    A = 0.25f * wsum * ((sig(-g, -g) + sig(g, -g)) + (sig(-g, g) + sig(g, g)));
}

// Nine series evaluations per pixel: the highest degree for every wave (it
// reaches every one-node pixel) keeps a single instance of the set-up.
__device__ inline bool sur_dispatch(const double* isep64, const PixelGeom<float>& pg, int r0,
                                    float wsum, float rx, float (&f)[4], float& A) {
    if (!__all(rx <= kPolyReach[kPolyMaxD - 2])) return false;
    sur_pixel<kPolyMaxD>(isep64, pg, r0, wsum, f, A);
    return true;
}

// Row weights of rows r0 .. r0+3 from the per-row sums S_k of w/4 and T_k
// of w t / 4 given as S0, T0, the cumulative S01, T01 and the totals wsum,
// Tt: row k receives S_k - T_k from its own points and T_{k-1} from those
// below it.
template <typename T>
__device__ inline void rows_from_sums(T wsum, T S0, T T0, T S01, T T01, T Tt, T (&W)[4]) {
    const T S1 = S01 - S0, T1 = T01 - T0, S2 = wsum - S01, T2 = Tt - T01;
    W[0] = S0 - T0;
    W[1] = T0 + (S1 - T1);
    W[2] = T1 + (S2 - T2);
    W[3] = T2;
}

// Pass 1.  Grid: (tile blocks, S); LDS: the in-radius row histogram
// [kNDPad] of double and, for FP64, 2 R2D / sep [kNDPad] and the rule (u,
// v, w/4, redm1 w/4) padded to a multiple of 4 points.  FP32 reads its rule
// arrays (LeanRule) and 2 R2D / sep from global memory.
//
// Every point of a pixel has its separation in [slo, shi], so the rows the
// pixel touches are known before the point loop.  A pixel whose points touch
// at most four consecutive rows r0 .. r0+3 accumulates, per relative row,
// the sums of w and w t in registers; the row weights follow from them.
// Other pixels are listed for near_rows_kernel.  When every such pixel of a
// wave lies below 28 degrees, the wave runs a branch-free FP32 loop over the
// points in groups of four (grid-structured rules reuse the blend of each
// rule row).  Occupancy: 6 waves per SIMD for FP32 (80 VGPRs), 5 for FP64.
//
// rowweight_block is one block's share: lin is the block's index on the
// (tile runs, sources) grid, or with PERSIST its index among nblk resident
// blocks; smem holds the dynamic LDS described above.
template <typename T, bool SUR = false, bool PERSIST = false>
__device__ inline void rowweight_block(const CubArgs& a, long long lin, long long nblk,
                                       unsigned char* smem) {
    using V4 = typename Vec4<T>::type;
    double* hist_s = reinterpret_cast<double*>(smem);
    T* isep_s = reinterpret_cast<T*>(hist_s + kNDPad);
    V4* rule_s = reinterpret_cast<V4*>(isep_s + kNDPad);
    const int N4 = (a.N + 3) & ~3;
    LeanRule lr;
#pragma unroll
    for (int j = 0; j < kLeanArrays; ++j) lr.p[j] = (cfloat*)(a.lean + j * N4);
    lr.isep64 = a.inv_sep64;
    __shared__ int sur_cnt_s[kThreads / 64];

    const int tid = threadIdx.x;
    const int ntiles = a.tiles_w * a.tiles_h;
    // The block's items in the flattened (source, tile) space: a run of
    // tiles_per_block tiles of one source on the (tile runs, sources) grid,
    // taken in tile-major order; with rw_persist (a 1-D grid of fewer, resident blocks) the
    // block's contiguous share of all S x ntiles items, taken as segments
    // split at source boundaries.
    long long i0, i1;
    int sx = 0;   // the source of a non-persistent block
    if (PERSIST) {
        const long long n = (long long)a.S * ntiles;
        i0 = n * lin / nblk;
        i1 = n * (lin + 1) / nblk;
    } else {
        // tile-major order of the (tile runs, sources) grid: consecutive
        // blocks take one tile run for consecutive sources, so the blocks in
        // flight share that run's pixel geometry in L2 (the whole image's,
        // 9 MB in FP64 at 400 x 400, would stream through it per source)
        sx = (int)(lin % a.S);
        const int tb = (int)(lin / a.S) * a.tiles_per_block;
        if (tb >= ntiles) return;
        i0 = (long long)sx * ntiles + tb;
        i1 = (long long)sx * ntiles + min(ntiles, tb + a.tiles_per_block);
    }
    if (i0 >= i1) return;

    if constexpr (sizeof(T) == 8) stage_rule<T>(a, isep_s, rule_s, N4, tid, kThreads);
    for (int k = tid; k < kND; k += kThreads) hist_s[k] = 0.0;
    __syncthreads();

  // one segment on the grid of tile runs (the loop runs once, and the
  // instance keeps no segment state across it)
  for (long long seg = i0; PERSIST ? seg < i1 : seg == i0;) {
    const int s = PERSIST ? (int)(seg / ntiles) : sx;
    const int t_begin = PERSIST ? (int)(seg % ntiles) : (int)(i0 - (long long)sx * ntiles);
    const int t_end = PERSIST ? (int)min((long long)ntiles, t_begin + (i1 - seg))
                              : min(ntiles, t_begin + a.tiles_per_block);
    seg = PERSIST ? seg + (t_end - t_begin) : i0 + 1;
    const SrcGeom g = a.geo[s];
    const bool do_hist = a.hist && g.in_fov;
    const int lx = tid % kTileW, ly = tid / kTileW;
    const int Wp1 = a.W + 1;
    const T j0 = (T)a.j0, jk = (T)a.jk, wsum = (T)a.wsum;
    const size_t plane = (size_t)a.H * a.W;
    V4* rw = reinterpret_cast<V4*>(a.rw) + (size_t)s * plane;
    int* rw_row = a.rw_row + (size_t)s * plane;
    int* near_list = a.near_list + (size_t)s * plane;

    for (int t = t_begin; t < t_end; ++t) {
        const int th = t / a.tiles_w, tw = t % a.tiles_w;
        const int h = th * kTileH + ly, w = tw * kTileW + lx;
        const bool valid = h < a.H && w < a.W;
        // out-of-image lanes compute an in-image pixel and store nothing, so
        // the point loops run in uniform control flow
        const int hc = min(h, a.H - 1), wc = min(w, a.W - 1);
        const PixelGeom<T> pg = pixel_geom<T>(a.corners, hc, wc, Wp1, g.dir, a.rho_scale);
        const int ilo = sep_index_fast(pg.slo), ihi = sep_index_fast(pg.shi);
        // rows touched: ilo-1 .. ihi; inside the table with t in [0, 1]
        const bool fast = (ihi - ilo) <= 2 && pg.slo >= T(1e-4) && pg.shi < T(69.9);
        const int r0 = min(max(ilo - 1, 0), kND - 4);

        // row weights of r0 .. r0+3; the FP64 and per-point paths form them
        // from S0 = sum_{rel=0} w, S01 = sum_{rel<=1} w, likewise T0, T01 of w t
        T W[4] = {T(0), T(0), T(0), T(0)};
        T S0 = T(0), T0 = T(0), S01 = T(0), T01 = T(0), Tt = T(0);
        // FP32: lanes that are not fast are listed for the near pass and
        // their sums are discarded, so they may run the lean loop (which
        // reads no memory that depends on the lane) on meaningless values.
        // FP64 indexes the table per point and keeps every lane in range.
        const bool lean6 = sizeof(T) == 4 ? __any(fast) && __all(!fast || pg.shi < T(28.0))
                                          : __all(fast && pg.shi < T(28.0));
        bool series = false;   // the wave's row weights came from a point loop below
        bool listed = false;   // handed to the surrogate pass (sur_rows_kernel)
        float sf[4] = {0.f, 0.f, 0.f, 0.f}, sA = 0.f;
        if constexpr (sizeof(T) == 4) {
            if (lean6) {
                // every fast pixel's points within three rows (one boundary)
                const bool one = __all(!fast || ihi - ilo <= 1);
                const float rx = fast ? poly_reach(pg) : 0.0f;
                // a fast pixel spans at most three row intervals, shi / slo <
                // e^(3 step), so rho < 0.051 and r < 0.104: degree 6 always
                // reaches; a wave it did not would take the per-point path
                if (SUR && one) {
                    // surrogate: A from four points; pixels whose bounds hold
                    // no node have B1 = |A| exactly, the others go to the
                    // network
                    series = sur_dispatch(lr.isep64, pg, r0, wsum, rx, sf, sA);
                    if (series) {
                        listed = fast && ihi == ilo + 1;
                        poly_rows<1>(wsum, sA, fabsf(sA), 0.0f, W);
                    }
                } else {
                    series = one ? poly_dispatch<1>(a, lr, pg, r0, N4, wsum, rx, W)
                                 : poly_dispatch<2>(a, lr, pg, r0, N4, wsum, rx, W);
                }
            }
        } else if (lean6) {
            series = true;
            // FP64: exact asin, row from the log of the separation
            for (int p0 = 0; p0 < N4; p0 += 4) {
                V4 r[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) r[k] = rule_s[p0 + k];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T u = r[k].x, v = r[k].y;
                    const T dx = pg.A[0] + u * pg.B[0] + v * (pg.C[0] + u * pg.E[0]);
                    const T dy = pg.A[1] + u * pg.B[1] + v * (pg.C[1] + u * pg.E[1]);
                    const T dz = pg.A[2] + u * pg.B[2] + v * (pg.C[2] + u * pg.E[2]);
                    far_point<T, 6>(dx, dy, dz, r[k].z, r[k].w, isep_s, j0, jk, r0, S0, T0, S01,
                                    T01, Tt);
                }
            }
            S01 += S0;
            T01 += T0;
            rows_from_sums(wsum, S0, T0, S01, T01, Tt, W);
        }
        if (!series && __any(fast)) {
            // ---- far pixels beyond 28 degrees or sharing a wave with near ones
            for (int p = 0; p < a.N; ++p) {
                T u, v, wq;
                if constexpr (sizeof(T) == 4) {
                    u = lr.p[0][p];
                    v = lr.p[1][p];
                    wq = lr.p[2][p];
                } else {
                    const V4 r = rule_s[p];
                    u = r.x;
                    v = r.y;
                    wq = r.z;
                }
                const T dx = pg.A[0] + u * pg.B[0] + v * (pg.C[0] + u * pg.E[0]);
                const T dy = pg.A[1] + u * pg.B[1] + v * (pg.C[1] + u * pg.E[1]);
                const T dz = pg.A[2] + u * pg.B[2] + v * (pg.C[2] + u * pg.E[2]);
                const T x = fast_sqrt(dx * dx + dy * dy + dz * dz);
                const T sd = T(2.0 * kR2D) * asin_pos(fmin(x, T(1)));
                const int i = sep_index_fast(sd);
                // t = (s - sep[i-1]) / (sep[i] - sep[i-1]) = redm1 (s / sep[i-1] - 1)
                const T is = sizeof(T) == 4 ? (T)a.inv_sep[i - 1] : isep_s[i - 1];
                const T tt = T(kRedm1) * (sd * (is * T(0.5 / kR2D)) - T(1));
                const int rel = min(max(i - 1 - r0, 0), 2);
                const T wt = wq * tt;
                S0 += rel == 0 ? wq : T(0);
                T0 += rel == 0 ? wt : T(0);
                S01 += rel <= 1 ? wq : T(0);
                T01 += rel <= 1 ? wt : T(0);
                Tt += wt;
            }
            rows_from_sums(wsum, S0, T0, S01, T01, Tt, W);
        }
        if constexpr (SUR) {
            // the tile's listed pixels, compacted in the tile's own 256 list
            // slots (no atomics): wave counts in LDS, then each lane's slot
            const bool app = valid && listed;
            const unsigned long long m = __ballot(app);
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            if (lane == 0) sur_cnt_s[wv] = __popcll(m);
            __syncthreads();
            int base = 0;
            for (int q = 0; q < wv; ++q) base += sur_cnt_s[q];
            if (app) {
                // the surrogate pass writes its row weights and histogram rows
                const int pix = h * a.W + w;
                rw_row[pix] = r0;
                const int k = base + __popcll(m & ((1ull << lane) - 1));
                const bool inr = do_hist && in_radius(g, a.centres, h * Wp1 + w);
                const float om = a.omega ? (float)a.omega[pix] : 1.0f;
                float4* e = a.sur_list + 2 * (((size_t)s * ntiles + t) * kThreads + k);
                e[0] = float4{sf[0], sf[1], sf[2], sf[3]};
                e[1] = float4{sA, __int_as_float(pix), om, __int_as_float(2 * r0 + (inr ? 1 : 0))};
            }
            if (threadIdx.x == 0) {
                int n = 0;
                for (int q = 0; q < kThreads / 64; ++q) n += sur_cnt_s[q];
                a.sur_count[(size_t)s * ntiles + t] = n;
            }
            __syncthreads();   // sur_cnt_s is rewritten by the next tile
        }
        if (valid && !listed) {
            const int pix = h * a.W + w;
            if (fast) {
                // row weights x solid angle
                const T om = a.omega ? (T)a.omega[pix] : T(1);
                const V4 wv = V4{om * W[0], om * W[1], om * W[2], om * W[3]};
                rw[pix] = wv;
                rw_row[pix] = r0;
                if (do_hist && in_radius(g, a.centres, h * Wp1 + w)) {
                    atomicAdd(&hist_s[r0], (double)wv.x);
                    atomicAdd(&hist_s[r0 + 1], (double)wv.y);
                    atomicAdd(&hist_s[r0 + 2], (double)wv.z);
                    atomicAdd(&hist_s[r0 + 3], (double)wv.w);
                }
            } else {
                rw_row[pix] = kRowListed;
                near_list[atomicAdd(&a.near_count[s], 1)] = pix;
            }
        }
    }
    if (do_hist) {
        // the segment's rows into the source's histogram (cleared for the
        // next segment of a persistent block)
        __syncthreads();
        for (int k = tid; k < kND; k += kThreads)
            if (hist_s[k] != 0.0) {
                atomicAdd(&a.hist[(size_t)s * kND + k], hist_s[k]);
                if (PERSIST) hist_s[k] = 0.0;
            }
        if (PERSIST) __syncthreads();
    }
  }
}

template <typename T, bool SUR = false, bool PERSIST = false>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(PERSIST ? 4 : (sizeof(T) == 4 ? 6 : 5))))
rowweight_kernel(CubArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if (PERSIST)
        rowweight_block<T, SUR, true>(a, blockIdx.x, gridDim.x, smem_raw);
    else
        rowweight_block<T, SUR, false>(a, blockIdx.x + (long long)gridDim.x * blockIdx.y, 0,
                                       smem_raw);
}


// Pass 1, a tile across many sources (FP32 cubature): a block takes one
// 16 x 16 tile and a range of sources.  The tile's source-independent pixel
// frame is formed once into LDS (half the corners' mean m, so that the
// blend's constant term is A = m - src / 2; the blend vectors B, C, E and the
// corners' spread hr0 of pixel_geom; the centre for the radius test; the
// solid angle), and so are the tables of 2 R2D / sep: the loop over the
// sources then reads no global memory but each source's SrcGeom (scalar
// loads, the next one's issued before the current one's work).  Beside a
// store-bound pass a load waits many microseconds; this pass issues none per
// source and keeps its compute busy.  The results equal rowweight_kernel's
// up to FP32 rounding (A, B, C, E are formed in FP64 from the corners and
// rounded once).  Items: (tile, run of sources), tile-major; block b takes
// items b, b + grid, ... (one per block, or a grid of a few resident blocks
// per CU that leaves room beside them for the store-bound emit).
struct PixFrame {
    double m[3];     // half the mean of the corners
    double c[3];     // centre
    float B[3], C[3], E[3];
    float hr0;       // max over corners |c_k / 2 - m|
    float om;        // solid angle (1 without)
};

__device__ inline bool in_radius_dir(const SrcGeom& g, const double* c) {
    const double ex = g.dir[0] - c[0], ey = g.dir[1] - c[1], ez = g.dir[2] - c[2];
    const double c2 = ex * ex + ey * ey + ez * ez;
    if (c2 < g.chord2_lo) return true;
    if (c2 > g.chord2_hi) return false;
    return in_radius_exact(c2, g.radius_deg);
}

// A source's geometry by scalar loads (the constant address space; written
// by an earlier launch, so the scalar cache holds no stale copy).
__device__ inline SrcGeom load_geo_scalar(const SrcGeom* p) {
    using cdouble = const __attribute__((address_space(4))) double;
    using cint = const __attribute__((address_space(4))) int;
    cdouble* d = (cdouble*)p;
    SrcGeom g;
    g.dir[0] = d[0];
    g.dir[1] = d[1];
    g.dir[2] = d[2];
    g.radius_deg = d[3];
    g.chord2_lo = d[4];
    g.chord2_hi = d[5];
    g.in_fov = ((cint*)p)[12];
    g.pad = 0;
    return g;
}

template <bool SUR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
rowweight_frames_kernel(CubArgs a, int nsplit) {
    __shared__ PixFrame fr_s[kThreads];
    __shared__ double isep_s[kNDPad];
    __shared__ double hist_s[kNDPad];
    __shared__ int sur_cnt_s[kThreads / 64];
    const int tid = threadIdx.x;
    const int ntiles = a.tiles_w * a.tiles_h;
    const long long nitem = (long long)ntiles * nsplit;
    if ((long long)blockIdx.x >= nitem) return;
    for (int r = tid; r < kNDPad; r += kThreads) {
        isep_s[r] = r < kND ? a.inv_sep64[r] : 0.0;
        hist_s[r] = 0.0;
    }
    const int N4 = (a.N + 3) & ~3;
    LeanRule lr;
#pragma unroll
    for (int j = 0; j < kLeanArrays; ++j) lr.p[j] = (cfloat*)(a.lean + j * N4);
    lr.isep64 = isep_s;
    const float wsum = (float)a.wsum;
    const float rho = (float)a.rho_scale;
    const size_t plane = (size_t)a.H * a.W;
    const int Wp1 = a.W + 1;
    const int lx = tid % kTileW, ly = tid / kTileW;
    __syncthreads();

  // items (tile item / nsplit, sources split nsplit ways) strided over the
  // grid: one item per block, or a few resident blocks per CU
  for (long long item = blockIdx.x; item < nitem; item += gridDim.x) {
    const int t = (int)(item / nsplit), ks = (int)(item % nsplit);
    const int s0 = (int)((long long)a.S * ks / nsplit);
    const int s1 = (int)((long long)a.S * (ks + 1) / nsplit);
    if (s0 >= s1) continue;
    const int th = t / a.tiles_w, tw = t % a.tiles_w;
    const int h = th * kTileH + ly, w = tw * kTileW + lx;
    const bool valid = h < a.H && w < a.W;
    // out-of-image lanes take an in-image pixel and store nothing
    const int hc = min(h, a.H - 1), wc = min(w, a.W - 1);
    const int pix = h * a.W + w;
    {
        // this thread's frame (only this thread reads its slot)
        const int c00 = hc * Wp1 + wc;
        const int cidx[4] = {c00, c00 + 1, c00 + Wp1 + 1, c00 + Wp1};  // LL LR UR UL
        double cc[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) cc[k][c] = a.corners[3 * cidx[k] + c];
        PixFrame f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            f.m[c] = 0.125 * ((cc[0][c] + cc[1][c]) + (cc[2][c] + cc[3][c]));
            f.B[c] = (float)(0.125 * ((cc[1][c] - cc[0][c]) + (cc[2][c] - cc[3][c])));
            f.C[c] = (float)(0.125 * ((cc[3][c] - cc[0][c]) + (cc[2][c] - cc[1][c])));
            f.E[c] = (float)(0.125 * ((cc[0][c] - cc[1][c]) + (cc[2][c] - cc[3][c])));
        }
        double hr = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double ex = 0.5 * cc[k][0] - f.m[0], ey = 0.5 * cc[k][1] - f.m[1],
                         ez = 0.5 * cc[k][2] - f.m[2];
            hr = fmax(hr, sqrt(ex * ex + ey * ey + ez * ez));
        }
        f.hr0 = (float)hr;
        const double* ce = a.centres + 3 * (size_t)c00;
        f.c[0] = ce[0];
        f.c[1] = ce[1];
        f.c[2] = ce[2];
        f.om = a.omega ? (float)a.omega[hc * a.W + wc] : 1.0f;
        fr_s[tid] = f;
    }

    // the sources' geometry by scalar loads (constant address space): vector
    // loads would wait behind the previous source's stores
    SrcGeom gn = load_geo_scalar(a.geo + s0);
    for (int s = s0; s < s1; ++s) {
        const SrcGeom g = gn;
        if (s + 1 < s1) gn = load_geo_scalar(a.geo + s + 1);
        const bool do_hist = a.hist && g.in_fov;
        const PixFrame f = fr_s[tid];
        PixelGeom<float> pg;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pg.A[c] = (float)(f.m[c] - 0.5 * g.dir[c]);
            pg.B[c] = f.B[c];
            pg.C[c] = f.C[c];
            pg.E[c] = f.E[c];
        }
        const float x0 = fast_sqrt(pg.A[0] * pg.A[0] + pg.A[1] * pg.A[1] + pg.A[2] * pg.A[2]);
        const float hr = f.hr0 * rho * 1.0001f + 1e-5f * x0;
        pg.x0 = x0;
        pg.hr = hr;
        pg.slo = float(2.0 * kR2D) * asin_pos(fmaxf(x0 - hr, 0.0f));
        pg.shi = float(2.0 * kR2D) * asin_pos(fminf(x0 + hr, 1.0f));
        const int ilo = sep_index_fast(pg.slo), ihi = sep_index_fast(pg.shi);
        const bool fast = (ihi - ilo) <= 2 && pg.slo >= 1e-4f && pg.shi < 69.9f;
        const int r0 = min(max(ilo - 1, 0), kND - 4);
        float W[4] = {0.f, 0.f, 0.f, 0.f};
        // lanes that are not fast are listed and their sums discarded (see
        // rowweight_block)
        const bool lean6 = __any(fast) && __all(!fast || pg.shi < 28.0f);
        bool series = false;
        bool listed = false;   // handed to the surrogate pass (sur_rows_kernel)
        float sf[4] = {0.f, 0.f, 0.f, 0.f}, sA = 0.f;
        if (lean6) {
            const bool one = __all(!fast || ihi - ilo <= 1);
            const float rx = fast ? poly_reach(pg) : 0.0f;
            if (SUR && one) {
                // as in rowweight_block: A from four points, B1 = |A| for
                // pixels whose bounds hold no node, the network for the others
                series = sur_dispatch(lr.isep64, pg, r0, wsum, rx, sf, sA);
                if (series) {
                    listed = fast && ihi == ilo + 1;
                    poly_rows<1>(wsum, sA, fabsf(sA), 0.0f, W);
                }
            } else {
                series = one ? poly_dispatch<1>(a, lr, pg, r0, N4, wsum, rx, W)
                             : poly_dispatch<2>(a, lr, pg, r0, N4, wsum, rx, W);
            }
        }
        if (!series && __any(fast)) {
            float S0 = 0.f, T0 = 0.f, S01 = 0.f, T01 = 0.f, Tt = 0.f;
            for (int p = 0; p < a.N; ++p) {
                const float u = lr.p[0][p], v = lr.p[1][p], wq = lr.p[2][p];
                const float dx = pg.A[0] + u * pg.B[0] + v * (pg.C[0] + u * pg.E[0]);
                const float dy = pg.A[1] + u * pg.B[1] + v * (pg.C[1] + u * pg.E[1]);
                const float dz = pg.A[2] + u * pg.B[2] + v * (pg.C[2] + u * pg.E[2]);
                const float x = fast_sqrt(dx * dx + dy * dy + dz * dz);
                const float sd = float(2.0 * kR2D) * asin_pos(fminf(x, 1.0f));
                const int i = sep_index_fast(sd);
                // (float)isep_s[k] is the FP32 table's entry
                const float tt =
                    float(kRedm1) * (sd * ((float)isep_s[i - 1] * float(0.5 / kR2D)) - 1.0f);
                const int rel = min(max(i - 1 - r0, 0), 2);
                const float wt = wq * tt;
                S0 += rel == 0 ? wq : 0.f;
                T0 += rel == 0 ? wt : 0.f;
                S01 += rel <= 1 ? wq : 0.f;
                T01 += rel <= 1 ? wt : 0.f;
                Tt += wt;
            }
            rows_from_sums(wsum, S0, T0, S01, T01, Tt, W);
        }
        if constexpr (SUR) {
            // the tile's listed pixels, compacted in its 256 list slots of
            // this source (wave counts in LDS, then each lane's slot)
            const bool app = valid && listed;
            const unsigned long long m = __ballot(app);
            const int lane = tid & 63, wv = tid >> 6;
            if (lane == 0) sur_cnt_s[wv] = __popcll(m);
            __syncthreads();
            int base = 0;
            for (int q = 0; q < wv; ++q) base += sur_cnt_s[q];
            if (app) {
                a.rw_row[(size_t)s * plane + pix] = r0;
                const int k = base + __popcll(m & ((1ull << lane) - 1));
                const bool inr = do_hist && in_radius_dir(g, f.c);
                float4* e = a.sur_list + 2 * (((size_t)s * ntiles + t) * kThreads + k);
                e[0] = float4{sf[0], sf[1], sf[2], sf[3]};
                e[1] = float4{sA, __int_as_float(pix), f.om,
                              __int_as_float(2 * r0 + (inr ? 1 : 0))};
            }
            if (tid == 0) {
                int n = 0;
                for (int q = 0; q < kThreads / 64; ++q) n += sur_cnt_s[q];
                a.sur_count[(size_t)s * ntiles + t] = n;
            }
            __syncthreads();   // sur_cnt_s is rewritten by the next source
        }
        if (valid && !listed) {
            if (fast) {
                const float4 wv = float4{f.om * W[0], f.om * W[1], f.om * W[2], f.om * W[3]};
                reinterpret_cast<float4*>(a.rw)[(size_t)s * plane + pix] = wv;
                a.rw_row[(size_t)s * plane + pix] = r0;
                if (do_hist && in_radius_dir(g, f.c)) {
                    atomicAdd(&hist_s[r0], (double)wv.x);
                    atomicAdd(&hist_s[r0 + 1], (double)wv.y);
                    atomicAdd(&hist_s[r0 + 2], (double)wv.z);
                    atomicAdd(&hist_s[r0 + 3], (double)wv.w);
                }
            } else {
                a.rw_row[(size_t)s * plane + pix] = kRowListed;
                a.near_list[(size_t)s * plane + atomicAdd(&a.near_count[s], 1)] = pix;
            }
        }
        if (do_hist) {
            __syncthreads();
            for (int r = tid; r < kND; r += kThreads) {
                const double v = hist_s[r];
                if (v != 0.0) {
                    atomicAdd(&a.hist[(size_t)s * kND + r], v);
                    hist_s[r] = 0.0;
                }
            }
            __syncthreads();
        }
    }
  }
}

// Pass 1b (surrogate mode): the listed one-node pixels of each source.  Each
// wave evaluates the network for 16 pixels at a time (xts_mlp.hpp), forms
// B1 = |A| + max(y, 0) (y = 0 without the network), the three row weights
// (poly_rows), stores them x solid angle and adds the in-radius ones to the
// block's LDS histogram, added to the source's histogram.  Grid: (blocks
// striding over the tiles, S), the histogram added once per block; or
// (resident != 0) a 1-D grid of resident blocks striding over (source, run
// of four tiles) items, the histogram added after each item (the pipelined
// map's second stream, beside the emit).  Each group's list entries are
// loaded while the previous group's network runs.  Rec: MlpPacked (BF16) or
// MlpPacked8 (FP8).
constexpr int kSurThreads = 256;
template <typename Rec>
__device__ inline void sur_tile(const CubArgs& a, const Rec* net, double* hblk, int s, int t,
                                int ntiles, bool use_net, int lane) {
    constexpr bool F8 = std::is_same_v<Rec, MlpPacked8>;
    const int count = a.sur_count[(size_t)s * ntiles + t];
    if (count <= 0) return;
    const float4* list = a.sur_list + 2 * ((size_t)s * ntiles + t) * kThreads;
    float4* rw = reinterpret_cast<float4*>(a.rw) + (size_t)s * a.H * a.W;
    const float wsum = (float)a.wsum;
    // the tile's entries in groups of 16 (the last group's spare lanes
    // evaluate the tile's last entry and store nothing)
    int ic = min(lane & 15, count - 1);
    float4 f = list[2 * ic], e = list[2 * ic + 1];
    for (int i0 = 0; i0 < count; i0 += 16) {
        const int i = i0 + (lane & 15);
        float4 fn = f, en = e;
        if (i0 + 16 < count) {
            ic = min(i + 16, count - 1);
            fn = list[2 * ic];
            en = list[2 * ic + 1];
        }
        float y = 0.0f;
        if (use_net) {
            const float x[4] = {f.x, f.y, f.z, f.w};
            if constexpr (F8) y = mlp_infer16_fp8(net, x);
            else y = mlp_infer16(net, x);
        }
        if (lane < 16 && i < count) {
            const float A = e.x, om = e.z;
            const int pix = __float_as_int(e.y), rf = __float_as_int(e.w);
            float W[4];
            poly_rows<1>(wsum, A, fabsf(A) + fmaxf(y, 0.0f), 0.0f, W);
            const float4 wv = float4{om * W[0], om * W[1], om * W[2], om * W[3]};
            rw[pix] = wv;
            if (rf & 1) {
                const int r0 = rf >> 1;
                atomicAdd(&hblk[r0], (double)wv.x);
                atomicAdd(&hblk[r0 + 1], (double)wv.y);
                atomicAdd(&hblk[r0 + 2], (double)wv.z);
            }
        }
        f = fn;
        e = en;
    }
}

template <typename Rec>
__global__ void __launch_bounds__(kSurThreads)
sur_rows_kernel(CubArgs a, int resident) {
    __shared__ Rec net;
    __shared__ double hblk[kNDPad];
    const int ntiles = a.tiles_w * a.tiles_h;
    const int waves = kSurThreads / 64, wave = threadIdx.x / 64, lane = threadIdx.x & 63;
    if (!resident && (int)blockIdx.x * waves >= ntiles) return;
    constexpr bool F8 = std::is_same_v<Rec, MlpPacked8>;
    const bool use_net = a.sur == (F8 ? 3 : 1);
    if (use_net) {
        const void* rec = F8 ? (const void*)a.sur_net8
                             : std::is_same_v<Rec, MlpPackedH> ? (const void*)a.sur_neth
                                                                : (const void*)a.sur_net;
        const float4* src = reinterpret_cast<const float4*>(rec);
        float4* dst = reinterpret_cast<float4*>(&net);
        for (int k = threadIdx.x; k < (int)(sizeof(Rec) / 16); k += kSurThreads)
            dst[k] = src[k];
    }
    for (int k = threadIdx.x; k < kND; k += kSurThreads) hblk[k] = 0.0;
    __syncthreads();
    if (resident) {
        const int64_t per_src = (ntiles + waves - 1) / waves;
        const int64_t nitem = (int64_t)a.S * per_src;
        for (int64_t it = blockIdx.x; it < nitem; it += gridDim.x) {
            const int s = (int)(it / per_src);
            const int t = (int)(it % per_src) * waves + wave;
            if (t < ntiles) sur_tile<Rec>(a, &net, hblk, s, t, ntiles, use_net, lane);
            __syncthreads();
            const bool do_hist = a.hist && a.geo[s].in_fov;
            for (int k = threadIdx.x; k < kND; k += kSurThreads) {
                const double v = hblk[k];
                if (v != 0.0) {
                    if (do_hist) atomicAdd(&a.hist[(size_t)s * kND + k], v);
                    hblk[k] = 0.0;
                }
            }
            __syncthreads();
        }
        return;
    }
    const int s = blockIdx.y;
    for (int t = blockIdx.x * waves + wave; t < ntiles; t += gridDim.x * waves)
        sur_tile<Rec>(a, &net, hblk, s, t, ntiles, use_net, lane);
    if (a.hist && a.geo[s].in_fov) {
        __syncthreads();
        for (int k = threadIdx.x; k < kND; k += kSurThreads)
            if (hblk[k] != 0.0) atomicAdd(&a.hist[(size_t)s * kND + k], hblk[k]);
    }
}

// Pass 2a (before the correction factor): the listed pixels of each source.
// A pixel whose points touch at most kNearRows consecutive rows rbase ..
// rbase+R-1 gathers its row weights and stores them, x solid angle, in a
// pool (R values at an offset in its source's segment); its in-radius
// weights go to the block's row sums in LDS, which are added to the
// source's row histogram once per block.  kNearSplit lanes of one wave
// share a pixel, each taking every kNearSplit-th rule point into a
// lane-private LDS histogram (rows x threads, no two lanes on one word); the
// lanes then sum the kNearSplit partial rows.  Below 14 degrees the FP32 row
// and fraction come from the relative log of lean_point.  A pixel touching
// more rows (the one holding the source), or finding the pool full, is
// "wide" (pool offset -1): only its in-radius weights are added here, point
// by point, and the emit's wide blocks evaluate it.  Grid: (blocks striding over
// the list, S).
constexpr int kNearRows = 24;
constexpr int kNearSplit = 2;   // 31.8 us against 34.7 with 4 (c1)
constexpr int kNearPix = kNearThreads / kNearSplit;   // pixels per block pass
template <typename T>
__global__ void __launch_bounds__(kNearThreads)
near_rows_kernel(CubArgs a) {
    using V4 = typename Vec4<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* isep_s = reinterpret_cast<T*>(smem_raw);
    V4* rule_s = reinterpret_cast<V4*>(isep_s + kNDPad);
    const int N4 = (a.N + 3) & ~3;
    float4* tco_s = reinterpret_cast<float4*>(rule_s + N4);
    T* ph = reinterpret_cast<T*>(tco_s + N4);   // [kNearRows][kNearThreads]
    // the block's in-radius row sums, added to the source's histogram once
    __shared__ double hblk[kNDPad];
    // the pass's wide in-radius pixels, whose rule points the block's waves
    // then take one pixel at a time (a lane per point, FP64)
    __shared__ int wide_s[kNearPix];
    __shared__ int wide_n;

    const int s = blockIdx.y;
    const int count = a.near_count[s];
    if ((int)blockIdx.x * kNearPix >= count) return;
    const int tid = threadIdx.x;
    const int slot = tid / kNearSplit, part = tid % kNearSplit;
    for (int k = tid; k < kND; k += kNearThreads) hblk[k] = 0.0;
    if (tid == 0) wide_n = 0;
    stage_rule<T>(a, isep_s, rule_s, N4, tid, kNearThreads);
    for (int p = tid; p < N4; p += kNearThreads) {
        const float wq = p < a.N ? (float)a.uvw[p].wq : 0.0f;
        tco_s[p] = float4{wq * (float)kTPoly[0], wq * (float)kTPoly[1], wq * (float)kTPoly[2], 0.0f};
    }
    __syncthreads();

    const SrcGeom g = a.geo[s];
    const bool do_hist = a.hist && g.in_fov;
    double* hist = do_hist ? a.hist + (size_t)s * kND : nullptr;
    const size_t plane = (size_t)a.H * a.W;
    const int Wp1 = a.W + 1;
    const int* list = a.near_list + (size_t)s * plane;
    int* meta_off = a.near_off + (size_t)s * plane;
    int* meta_rr = a.near_rr + (size_t)s * plane;
    T* pool = reinterpret_cast<T*>(a.near_pool);
    T* myh = ph + tid;                              // row r at myh[r * kNearThreads]
    for (int first = blockIdx.x * kNearPix; first < count; first += gridDim.x * kNearPix) {
        const int idx = first + slot;
        const bool live = idx < count;
        const int pix = live ? list[idx] : list[first];
        const int h = pix / a.W, w = pix % a.W;
        const PixelGeom<T> pg = pixel_geom<T>(a.corners, h, w, Wp1, g.dir, a.rho_scale);
        const T om = a.omega ? (T)a.omega[pix] : T(1);
        const bool inrad = live && do_hist && in_radius(g, a.centres, h * Wp1 + w);
        // rows touched: rbase = lower row of slo .. the upper row of shi
        const int rbase = sep_index_fast(pg.slo) - 1;
        const int R = sep_index_fast(pg.shi) - rbase + 1;
        // a fixed slot of kNearRows entries per list position in this
        // source's segment of the pool (no atomic in the pixel's chain)
        // (idx < cap / kNearRows: no product that could overflow; s * cap and
        // idx * kNearRows stay below 2^30)
        const int off = live && R <= kNearRows && idx < a.near_pool_cap / kNearRows
                            ? s * a.near_pool_cap + idx * kNearRows
                            : -1;
        const bool narrow = off >= 0;
        if (narrow)
            for (int r = 0; r < R; ++r) myh[r * kNearThreads] = T(0);
        bool lean = false;
        if constexpr (sizeof(T) == 4) lean = narrow && pg.slo >= T(1e-4) && pg.shi < T(14.0);
        if (lean) {
            if constexpr (sizeof(T) == 4) {
                const LeanPixel lp = lean_pixel(pg, rbase);
#pragma unroll 4
                for (int p = part; p < a.N; p += kNearSplit) {
                    const V4 rr = rule_s[p];
                    const float4 tc = tco_s[p];
                    const float u = rr.x, v = rr.y;
                    const float au = fmaf(fmaf(lp.a2, u, lp.a1), u, lp.a0);
                    const float bu = fmaf(fmaf(lp.b2, u, lp.b1), u, lp.b0);
                    const float cu = fmaf(fmaf(lp.c2, u, lp.c1), u, lp.c0);
                    const float z = fmaf(fmaf(cu, v, bu), v, au);
                    const float L = __builtin_amdgcn_logf(z);
                    const float gq = fmaf(fmaf(lp.g[2], z, lp.g[1]), z, lp.g[0]);
                    const float Q = fmaf(gq, z, fmaf((float)(0.5 * kRecipStep * 0.69314718055994530942),
                                                    L, lp.q0));
                    const float f = __builtin_amdgcn_fractf(Q);
                    const float wt = f * fmaf(f, fmaf(f, tc.z, tc.y), tc.x);
                    const int l = min(max((int)floorf(Q), 0), R - 2);
                    myh[l * kNearThreads] += (T)(rr.z - wt);
                    myh[(l + 1) * kNearThreads] += (T)wt;
                }
            }
        } else if (!narrow) {
            // a wide pixel's in-radius weights: after this pass, by a wave
            if (inrad && part == 0) wide_s[atomicAdd(&wide_n, 1)] = pix;
        } else {
            for (int p = part; p < a.N; p += kNearSplit) {
                const V4 rr = rule_s[p];
                const T u = rr.x, v = rr.y, wq = rr.z;
                const T dx = pg.A[0] + u * pg.B[0] + v * (pg.C[0] + u * pg.E[0]);
                const T dy = pg.A[1] + u * pg.B[1] + v * (pg.C[1] + u * pg.E[1]);
                const T dz = pg.A[2] + u * pg.B[2] + v * (pg.C[2] + u * pg.E[2]);
                const T x = fast_sqrt(dx * dx + dy * dy + dz * dz);
                const T sd = T(2.0 * kR2D) * asin_pos(fmin(x, T(1)));
                const int i = sep_index_fast(sd);
                const T tt = T(kRedm1) * (sd * (isep_s[i - 1] * T(0.5 / kR2D)) - T(1));
                const T wa = wq * (T(1) - tt), wb = wq * tt;
                const int l = min(max(i - 1 - rbase, 0), R - 2);
                myh[l * kNearThreads] += wa;
                myh[(l + 1) * kNearThreads] += wb;
            }
        }
        if (narrow) {
            // the pixel's kNearSplit partial histograms are in this wave's
            // adjacent lanes tid - part .. tid - part + kNearSplit - 1
            T* base = ph + (tid - part);
            for (int r = part; r < R; r += kNearSplit) {
                T W = T(0);
#pragma unroll
                for (int q = 0; q < kNearSplit; ++q) W += base[r * kNearThreads + q];
                W *= om;
                pool[off + r] = W;
                if (inrad) atomicAdd(&hblk[rbase + r], (double)W);
            }
        }
        if (live && part == 0) {
            meta_off[idx] = off;
            meta_rr[idx] = rbase * 256 + min(R, 255);
            if (off < 0) {
                a.rw_row[(size_t)s * plane + pix] = kRowWide;
                a.wide_list[(size_t)s * plane + atomicAdd(&a.wide_count[s], 1)] = idx;
            }
        }
        // the pass's wide pixels: a lane per rule point, the points' offsets
        // from the source in FP64 (they cancel the pixel-sized blend terms;
        // see wide_pixels)
        __syncthreads();
        const int nw = wide_n;
        const int lane = tid & 63;
        for (int k = tid >> 6; k < nw; k += kNearThreads / 64) {
            const int wp = wide_s[k];
            const double omd = a.omega ? a.omega[wp] : 1.0;
            const PixelGeom<double> pd =
                pixel_geom<double>(a.corners, wp / a.W, wp % a.W, Wp1, g.dir, a.rho_scale);
            for (int p = lane; p < a.N; p += 64) {
                const RuleUVW q = a.uvw[p];
                const double u = q.u, v = q.v;
                const double dx = pd.A[0] + u * pd.B[0] + v * (pd.C[0] + u * pd.E[0]);
                const double dy = pd.A[1] + u * pd.B[1] + v * (pd.C[1] + u * pd.E[1]);
                const double dz = pd.A[2] + u * pd.B[2] + v * (pd.C[2] + u * pd.E[2]);
                const double sd =
                    2.0 * kR2D * asin(fmin(sqrt(dx * dx + dy * dy + dz * dz), 1.0));
                const int i = sep_index_fast(sd);
                const double tt = kRedm1 * (sd * (a.inv_sep64[i - 1] * (0.5 / kR2D)) - 1.0);
                atomicAdd(&hblk[i - 1], omd * (q.wq * (1.0 - tt)));
                atomicAdd(&hblk[i], omd * (q.wq * tt));
            }
        }
        // every wave is past the list before it is cleared, and the clear is
        // seen before the next pass adds to it
        __syncthreads();
        if (tid == 0) wide_n = 0;
        __syncthreads();
    }
    if (do_hist) {
        __syncthreads();
        for (int k = tid; k < kND; k += kNearThreads)
            if (hblk[k] != 0.0) atomicAdd(&hist[k], hblk[k]);
    }
}

// Wide listed pixels (pool offset -1): one wave per pixel, run by extra
// blocks of the emit kernel (blockIdx.y = source, blockIdx.z = energy
// chunk).  The k-th entry of the source's wide list goes to wave
// gw = k mod nwaves.  The wave's lanes take the rule points, adding
// each point's two row weights to a 401-row LDS histogram H (the wave's
// own); then each lane takes every 64th table row, loads its EC energies as
// vectors and accumulates W_r L[r], and the wave sums the 64 partial
// vectors.  x solid angle, x exposure x correction when SCALED.
constexpr int kWideBlocks = 2;   // extra emit blocks per source and energy chunk
template <typename T, int EC, typename OT, bool SCALED, bool SUM = false>
__device__ inline void wide_pixels(const CubArgs& a, int gw, int nwaves, T* H, int s, int e0) {
    using V = typename Vec<T>::type;
    constexpr int VN = Vec<T>::n;
    constexpr int NV = EC / VN;
    const int lane = threadIdx.x % 64;
    const int ne = min(EC, a.E - e0);
    const SrcGeom g = a.geo[s];
    const size_t plane = (size_t)a.H * a.W;
    const int* list = a.near_list + (size_t)s * plane;
    const T* lut_s0 = reinterpret_cast<const T*>(a.lut) + (size_t)s * kND * a.Ep + e0;
    // this wave's entries of the source's wide list (written by the near pass)
    const int nw = a.wide_count[s];
    for (int k = gw; k < nw; k += nwaves) {
        const int pix = list[a.wide_list[(size_t)s * plane + k]];
        const int h = pix / a.W, w = pix % a.W;
        // The points' offsets from the source in FP64 whatever T: a wide
        // pixel holds (or touches) the source, and its points' offsets cancel
        // the pixel-sized blend terms to a small fraction of them.  FP32
        // would leave errors of 1e-4 in t that a table varying fast from row
        // to row turns into errors of 1e-4 in the value.
        const PixelGeom<double> pg =
            pixel_geom<double>(a.corners, h, w, a.W + 1, g.dir, a.rho_scale);
        const T om = a.omega ? (T)a.omega[pix] : T(1);
        for (int r = lane; r < kND; r += 64) H[r] = T(0);
        for (int p = lane; p < a.N; p += 64) {
            const RuleUVW q = a.uvw[p];
            const double u = q.u, v = q.v;
            const double dx = pg.A[0] + u * pg.B[0] + v * (pg.C[0] + u * pg.E[0]);
            const double dy = pg.A[1] + u * pg.B[1] + v * (pg.C[1] + u * pg.E[1]);
            const double dz = pg.A[2] + u * pg.B[2] + v * (pg.C[2] + u * pg.E[2]);
            const double x = sqrt(dx * dx + dy * dy + dz * dz);
            const double sd = 2.0 * kR2D * asin(fmin(x, 1.0));
            const int i = sep_index_fast(sd);
            const double tt = kRedm1 * (sd * (a.inv_sep64[i - 1] * (0.5 / kR2D)) - 1.0);
            atomicAdd(&H[i - 1], (T)(q.wq * (1.0 - tt)));
            atomicAdd(&H[i], (T)(q.wq * tt));
        }
        V acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = V{};
        for (int r = lane; r < kND; r += 64) {
            const T W = H[r];
            if (W != T(0)) {
                const V* L = reinterpret_cast<const V*>(lut_s0 + (size_t)r * a.Ep);
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    if (e0 + v * VN < a.Ep) vfma(acc[v], W, L[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int j = 0; j < VN; ++j) {
                T x = reinterpret_cast<const T*>(&acc[v])[j];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
                const int e = v * VN + j;
                if (lane == 0 && e < ne) {
                    x *= om;
                    if constexpr (SUM) {
                        // summed model: x flux into the [E][H W] sum
                        const int se = s * a.E + e0 + e;
                        atomicAdd(a.sum + (size_t)(e0 + e) * plane + pix,
                                  (float)x * (float)(a.expo[se] * a.cf[se] * a.flux[s]));
                        continue;
                    }
                    OT* po = reinterpret_cast<OT*>(a.out) +
                             out_index(a.layout_sehw, s, pix, e0 + e, a.E, plane);
                    if (SCALED) {
                        const int se = s * a.E + e0 + e;
                        if constexpr (sizeof(OT) == 4) {
                            *po = (float)x * (float)(a.expo[se] * a.cf[se]);
                        } else {
                            double y = (double)x * a.expo[se];
                            *po = y * a.cf[se];
                        }
                    } else {
                        *po = (OT)x;
                    }
                }
            }
        }
    }
}

// Pass 2b (after the correction factor): the pooled listed pixels' map
// values, written once: sum_r W_r L[rbase + r] (W already x solid angle),
// x exposure x correction when SCALED.  Grid: (blocks striding over the
// list, S, energy chunks).  SUM: added x flux to a summed model instead.
template <typename T, int EC, typename OT, bool SCALED, bool SUM = false>
__global__ void __launch_bounds__(kNearThreads)
near_emit_kernel(CubArgs a) {
    using V = typename Vec<T>::type;
    constexpr int VN = Vec<T>::n;
    constexpr int NV = EC / VN;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lut_s = reinterpret_cast<T*>(smem_raw);
    __shared__ ChunkScale<EC> sc;

    const int s = blockIdx.y;
    const int e0 = blockIdx.z * EC;
    const int ne = min(EC, a.E - e0);
    const int count = a.near_count[s];
    if ((int)blockIdx.x * kNearThreads >= count) return;
    const int tid = threadIdx.x;
    stage_lut<T, EC, kNearThreads>(a, s, e0, lut_s, tid);
    if (SCALED) sc.load(a, s * a.E + e0, ne, tid);
    __syncthreads();

    const size_t plane = (size_t)a.H * a.W;
    const int* list = a.near_list + (size_t)s * plane;
    const int* meta_off = a.near_off + (size_t)s * plane;
    const int* meta_rr = a.near_rr + (size_t)s * plane;
    const T* pool = reinterpret_cast<const T*>(a.near_pool);
    for (int first = blockIdx.x * kNearThreads; first < count;
         first += gridDim.x * kNearThreads) {
        const int idx = first + tid;
        const bool live = idx < count;
        const int ic = live ? idx : first;
        const int pix = list[ic];
        const int off = meta_off[ic];
        V acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = V{};
        if (off < 0) continue;            // wide: the emit's wide blocks
        {
            const int rr = meta_rr[ic];
            const int rbase = rr >> 8, R = rr & 255;
            const V* L = reinterpret_cast<const V*>(lut_s) + rbase * NV;
            for (int r0 = 0; r0 < R; r0 += 8) {
                // eight pool loads in flight, then their rows
                T Wb[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) Wb[k] = r0 + k < R ? pool[off + r0 + k] : T(0);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (r0 + k < R) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) vfma(acc[v], Wb[k], L[(r0 + k) * NV + v]);
                    }
                }
            }
        }
        if (!live) continue;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int j = 0; j < VN; ++j) {
                const int e = v * VN + j;
                if (e < ne) {
                    const T x = reinterpret_cast<const T*>(&acc[v])[j];
                    if constexpr (SUM) {
                        atomicAdd(a.sum + (size_t)(e0 + e) * plane + pix,
                                  sc.template apply<float>(x, e));
                    } else {
                        OT* po = reinterpret_cast<OT*>(a.out) +
                                 out_index(a.layout_sehw, s, pix, e0 + e, a.E, plane);
                        *po = SCALED ? sc.template apply<OT>(x, e) : (OT)x;
                    }
                }
            }
        }
    }
}

// Pass 3: per (source, pixel range, energy chunk): map = sum_k W_k L[r0+k]
// for pass-1 pixels, x exposure x correction when SCALED.  Listed pixels are
// left to near_emit_kernel, wide ones to WIDE's extra blocks or to
// wide_emit_kernel (a launch of its own: its registers would limit the
// emit's residency beside other work, as in the pipelined map).  OT: output type.  Each thread first loads the
// row weights of its kEmitPix pixels and then only stores: a load issued
// after stores would wait for them to drain (the vector memory counter
// counts both).
//
// [S, E, H, W]: the energy loop is outermost, so a block writes runs of
// kEmitThreads consecutive values of one plane; the table rows are staged
// with an odd stride (EC + 1) so lanes on different rows read different LDS
// banks.  [S, H, W, E]: each thread writes its pixels' energies as vectors.
constexpr int kEmitPix = 4;
constexpr int kEmitBlockPix = kEmitThreads * kEmitPix;
constexpr int kEmitEC32 = 40, kEmitEC64 = 16;   // energies per block (LDS: 401 rows of EC + 1)
template <typename T, int EC, typename OT, bool SCALED, bool WIDE>
__global__ void __launch_bounds__(kEmitThreads)
emit_sehw_kernel(CubArgs a) {
    using V4 = typename Vec4<T>::type;
    constexpr int LS = EC + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lut_s = reinterpret_cast<T*>(smem_raw);
    __shared__ ChunkScale<EC> sc;
    if (WIDE && (int)blockIdx.x >= a.emit_gx) {
        constexpr int kWaves = kEmitThreads / 64;
        wide_pixels<T, EC, OT, SCALED>(a, (blockIdx.x - a.emit_gx) * kWaves + threadIdx.x / 64,
                                       kWideBlocks * kWaves, lut_s + (threadIdx.x / 64) * kNDPad,
                                       blockIdx.y, blockIdx.z * EC);
        return;
    }

    // beside a compute-bound pass, the stores' waves may take issue first
    if (a.emit_prio == 1) __builtin_amdgcn_s_setprio(1);
    else if (a.emit_prio == 2) __builtin_amdgcn_s_setprio(2);
    else if (a.emit_prio >= 3) __builtin_amdgcn_s_setprio(3);
    const int s = blockIdx.y;
    const int e0 = blockIdx.z * EC;
    const int ne = min(EC, a.E - e0);
    const int tid = threadIdx.x;
    const int plane = a.H * a.W;
    const int se0 = s * a.E + e0;
    const V4* rw = reinterpret_cast<const V4*>(a.rw) + (size_t)s * plane;
    const int* rw_row = a.rw_row + (size_t)s * plane;
    // planes of a multiple of 4 pixels: each thread takes 4 consecutive
    // pixels and stores them as one vector per energy
    const bool vec = (plane & 3) == 0;
    int pix[kEmitPix], lrow[kEmitPix];
    bool out_ok[kEmitPix];
    V4 wv[kEmitPix];
    int r0[kEmitPix];
#pragma unroll
    for (int k = 0; k < kEmitPix; ++k) {
        pix[k] = blockIdx.x * kEmitBlockPix + (vec ? 4 * tid + k : k * kEmitThreads + tid);
        // loads from a clamped index, unconditionally, so they overlap
        const int pc = min(pix[k], plane - 1);
        r0[k] = rw_row[pc];
        wv[k] = rw[pc];
    }
    // Pixels of the near pool are written too, as zeros, and overwritten by
    // near_emit_kernel (later on the same stream): whole 128-byte lines
    // instead of lines with holes around each source, 145 -> 119 us for
    // 100 sources x 37 planes of 200 x 200.  Wide pixels are left to the
    // wide blocks of this kernel.
#pragma unroll
    for (int k = 0; k < kEmitPix; ++k) {
        out_ok[k] = pix[k] < plane && r0[k] != kRowWide;
        const bool far = out_ok[k] && r0[k] >= 0;
        lrow[k] = far ? r0[k] * LS : 0;
        if (!far) wv[k] = V4{T(0), T(0), T(0), T(0)};
    }
    const bool all_ok = out_ok[0] && out_ok[1] && out_ok[2] && out_ok[3];
    OT* po = reinterpret_cast<OT*>(a.out) + (size_t)se0 * plane;
    stage_lut<T, EC, kEmitThreads, LS>(a, s, e0, lut_s, tid);
    if (SCALED) sc.load(a, se0, ne, tid);
    __syncthreads();

#pragma unroll 2
    for (int e = 0; e < ne; ++e, po += plane) {
        OT y[kEmitPix];
#pragma unroll
        for (int k = 0; k < kEmitPix; ++k) {
            const T* L = lut_s + lrow[k] + e;
            T x = wv[k].x * L[0];
            x = fma(wv[k].y, L[LS], x);
            x = fma(wv[k].z, L[2 * LS], x);
            x = fma(wv[k].w, L[3 * LS], x);
            y[k] = SCALED ? sc.template apply<OT>(x, e) : (OT)x;
        }
        if (vec && all_ok) {
            stream_store4(po + pix[0], y[0], y[1], y[2], y[3]);
        } else {
#pragma unroll
            for (int k = 0; k < kEmitPix; ++k)
                if (out_ok[k]) stream_store(po + pix[k], y[k]);
        }
    }
}

template <typename T, int EC, typename OT, bool SCALED, bool WIDE>
__global__ void __launch_bounds__(kEmitThreads)
emit_shwe_kernel(CubArgs a) {
    using V = typename Vec<T>::type;
    constexpr int VN = Vec<T>::n;
    constexpr int NV = EC / VN;
    using V4 = typename Vec4<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lut_s = reinterpret_cast<T*>(smem_raw);
    __shared__ ChunkScale<EC> sc;
    if (WIDE && (int)blockIdx.x >= a.emit_gx) {
        constexpr int kWaves = kEmitThreads / 64;
        wide_pixels<T, EC, OT, SCALED>(a, (blockIdx.x - a.emit_gx) * kWaves + threadIdx.x / 64,
                                       kWideBlocks * kWaves, lut_s + (threadIdx.x / 64) * kNDPad,
                                       blockIdx.y, blockIdx.z * EC);
        return;
    }

    const int s = blockIdx.y;
    const int e0 = blockIdx.z * EC;
    const int ne = min(EC, a.E - e0);
    const int tid = threadIdx.x;
    const int plane = a.H * a.W;
    const int base = blockIdx.x * kEmitBlockPix;
    const int se0 = s * a.E + e0;
    const V4* rw = reinterpret_cast<const V4*>(a.rw) + (size_t)s * plane;
    const int* rw_row = a.rw_row + (size_t)s * plane;
    int pix[kEmitPix], r0[kEmitPix];
    V4 wv[kEmitPix];
#pragma unroll
    for (int k = 0; k < kEmitPix; ++k) {
        pix[k] = base + k * kEmitThreads + tid;
        const int pc = min(pix[k], plane - 1);
        r0[k] = rw_row[pc];
        wv[k] = rw[pc];
    }
#pragma unroll
    for (int k = 0; k < kEmitPix; ++k) {
        if (pix[k] >= plane) r0[k] = kRowWide;   // nothing to write
        // pool pixels as zeros (near_emit_kernel writes them later)
        if (r0[k] == kRowListed) wv[k] = V4{T(0), T(0), T(0), T(0)};
    }
    stage_lut<T, EC, kEmitThreads>(a, s, e0, lut_s, tid);
    if (SCALED) sc.load(a, se0, ne, tid);
    __syncthreads();

#pragma unroll
    for (int k = 0; k < kEmitPix; ++k) {
        if (r0[k] == kRowWide) continue;
        const V* L = reinterpret_cast<const V*>(lut_s) + max(r0[k], 0) * NV;
        OT* po = reinterpret_cast<OT*>(a.out) + ((size_t)s * plane + pix[k]) * a.E + e0;
#pragma unroll 2
        for (int v = 0; v < NV; ++v) {
            V acc = V{};
            vfma(acc, wv[k].x, L[v]);
            vfma(acc, wv[k].y, L[NV + v]);
            vfma(acc, wv[k].z, L[2 * NV + v]);
            vfma(acc, wv[k].w, L[3 * NV + v]);
#pragma unroll
            for (int j = 0; j < VN; ++j) {
                const int e = v * VN + j;
                if (e < ne) {
                    const T x = reinterpret_cast<const T*>(&acc)[j];
                    stream_store(po + e, SCALED ? sc.template apply<OT>(x, e) : (OT)x);
                }
            }
        }
    }
}


// The summed model's per-(source, energy) factor flux x exposure x
// correction, rounded to FP32 once ([S][Ep], zero past E).
__global__ void sum_factor_kernel(CubArgs a) {
    const int n = a.S * a.Ep;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int s = k / a.Ep, e = k % a.Ep;
        a.sum_fac[k] = e < a.E ? (float)(a.expo[s * a.E + e] * a.cf[s * a.E + e] * a.flux[s])
                               : 0.0f;
    }
}

// Zeroes n floats (a kernel rather than a memset, so that captured graphs
// replay it correctly).
__global__ void zero_kernel(float* __restrict__ p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0f;
}

// Pass 3 of a summed model (xts_point_src_summed_model_dev), FP32 on the
// matrix cores: per 16x16 pixel tile and chunk of 16 NT energies, the sum
// over the batch's sources of flux x exposure x correction x sum_k W_k
// L[r0 + k] for the far pixels.  The block takes the sources in groups of
// G: it loads the group's row indices and weights together, stages each
// source's scaled table rows M[r][e] = L[r][e] x factor[e] over its far
// pixels' window [rmin, rmax + 3] (at most kSumWin rows) in LDS, and then
// each wave forms D[pixel][e] += A[pixel][r] M[r][e] on
// v_mfma_f32_16x16x4_f32 for every source of the group, four window rows a
// step: A holds each pixel's four row weights in its rows' columns (zero
// elsewhere).  One load latency and two barriers serve G sources.  Wave w
// owns tile rows 4w .. 4w+3 (one 16-pixel m-tile each); lane l owns the
// pixel of m-tile l / 16, column l % 16, and supplies A[l % 16][l / 16] for
// each m-tile from that pixel's lane by a shuffle; B[k][n] = M[4 kk + l /
// 16][16 n + l % 16]; D[4 (l / 16) + i][l % 16] is four consecutive pixels
// of one energy.  Wider windows (tiles next to the source) run afterwards,
// one source at a time, in passes of G kSumWin rows.  Listed and wide
// pixels are added by near_emit_kernel and wide_sum_kernel; the launches of
// one call are ordered on one stream, so the read-add-write of the sum
// needs no atomics.
constexpr int kSumWin = 16;
typedef float sum_f32x4 __attribute__((ext_vector_type(4)));

template <int NT>
__device__ inline void sum_mfma_window(const float* __restrict__ Mw, int c0, int nk, int lane,
                                       const int (&ra)[4], const float (&wa)[4][4],
                                       sum_f32x4 (&acc)[4][NT]) {
    constexpr int EC = 16 * NT;
    for (int kk = 0; kk < nk; ++kk) {
        // A[l % 16][l / 16]: the weight of window row c0 + 4 kk + l / 16
        const int row = c0 + 4 * kk + (lane >> 4);
        float A[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int j = row - ra[m];
            A[m] = j == 0 ? wa[m][0] : j == 1 ? wa[m][1] : j == 2 ? wa[m][2]
                                                          : j == 3 ? wa[m][3] : 0.0f;
        }
        const float* Mk = Mw + (4 * kk + (lane >> 4)) * EC + (lane & 15);
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const float B = Mk[16 * n];
#pragma unroll
            for (int m = 0; m < 4; ++m)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[m], B, acc[m][n], 0, 0, 0);
        }
    }
}

// Rows [c0, c0 + nr) of source s's scaled table into Mw (4 nk rows, the
// rest zero), by all threads of the block.
template <int EC>
__device__ inline void sum_stage(const CubArgs& a, int s, int e0, int ne, int c0, int nr, int nk,
                                 float* __restrict__ Mw, int tid) {
    const float* L = reinterpret_cast<const float*>(a.lut) + (size_t)s * kND * a.Ep + e0;
    const float* f = a.sum_fac + (size_t)s * a.Ep + e0;
    for (int k = tid; k < 4 * nk * EC; k += kThreads) {
        const int r = k / EC, e = k % EC;
        float x = 0.0f;
        if (r < nr && e < ne) x = L[(size_t)(c0 + r) * a.Ep + e] * f[e];
        Mw[k] = x;
    }
}

template <int NT, int G>
__global__ void __launch_bounds__(kThreads)
sum_far_kernel(CubArgs a) {
    constexpr int EC = 16 * NT;
    constexpr int kWaves = kThreads / 64;
    __shared__ __align__(16) float M[G * kSumWin * EC];
    __shared__ int red[G][2 * kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int th = blockIdx.x / a.tiles_w, tw = blockIdx.x % a.tiles_w;
    // this lane's pixel: m-tile lane / 16 of the wave, column lane % 16
    const int h = th * kTileH + 4 * wv + (lane >> 4), w = tw * kTileW + (lane & 15);
    const bool valid = h < a.H && w < a.W;
    const int pix = valid ? h * a.W + w : 0;
    const int e0 = blockIdx.y * EC;
    const int ne = min(EC, a.E - e0);
    const size_t plane = (size_t)a.H * a.W;
    const float4* rw4 = reinterpret_cast<const float4*>(a.rw);
    sum_f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = sum_f32x4{0.f, 0.f, 0.f, 0.f};
    for (int s0 = 0; s0 < a.S; s0 += G) {
        const int ng = min(G, a.S - s0);
        int r0g[G];
        float4 Wg[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            r0g[g] = kRowWide;
            Wg[g] = float4{0.f, 0.f, 0.f, 0.f};
            if (g < ng && valid) {
                r0g[g] = a.rw_row[(size_t)(s0 + g) * plane + pix];
                Wg[g] = rw4[(size_t)(s0 + g) * plane + pix];
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const bool far = r0g[g] >= 0;
            int lo = far ? r0g[g] : kND, hi = far ? r0g[g] : -1;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                lo = min(lo, __shfl_xor(lo, o));
                hi = max(hi, __shfl_xor(hi, o));
            }
            if (lane == 0) {
                red[g][wv] = lo;
                red[g][kWaves + wv] = hi;
            }
        }
        __syncthreads();
        int rmin[G], rmax[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            rmin[g] = kND;
            rmax[g] = -1;
#pragma unroll
            for (int q = 0; q < kWaves; ++q) {
                rmin[g] = min(rmin[g], red[g][q]);
                rmax[g] = max(rmax[g], red[g][kWaves + q]);
            }
            if (g >= ng) rmax[g] = -1;
        }
        // the group's narrow windows, staged together
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nr = rmax[g] + 4 - rmin[g];
            if (rmax[g] >= 0 && nr <= kSumWin)
                sum_stage<EC>(a, s0 + g, e0, ne, rmin[g], nr, (nr + 3) / 4,
                              M + g * kSumWin * EC, tid);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nr = rmax[g] + 4 - rmin[g];
            if (rmax[g] < 0 || nr > kSumWin) continue;
            const int rr = r0g[g] >= 0 ? r0g[g] : -1000;   // not far: out of every window
            int ra[4];
            float wa[4][4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int src = 16 * m + (lane & 15);
                ra[m] = __shfl(rr, src);
                wa[m][0] = __shfl(Wg[g].x, src);
                wa[m][1] = __shfl(Wg[g].y, src);
                wa[m][2] = __shfl(Wg[g].z, src);
                wa[m][3] = __shfl(Wg[g].w, src);
            }
            sum_mfma_window<NT>(M + g * kSumWin * EC, rmin[g], (nr + 3) / 4, lane, ra, wa, acc);
        }
        // wide windows: one source at a time, in passes of G kSumWin rows
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (rmax[g] < 0 || rmax[g] + 4 - rmin[g] <= kSumWin) continue;
            const int rr = r0g[g] >= 0 ? r0g[g] : -1000;
            int ra[4];
            float wa[4][4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int src = 16 * m + (lane & 15);
                ra[m] = __shfl(rr, src);
                wa[m][0] = __shfl(Wg[g].x, src);
                wa[m][1] = __shfl(Wg[g].y, src);
                wa[m][2] = __shfl(Wg[g].z, src);
                wa[m][3] = __shfl(Wg[g].w, src);
            }
            for (int c0 = rmin[g]; c0 <= rmax[g] + 3; c0 += G * kSumWin) {
                const int nr = min(G * kSumWin, rmax[g] + 4 - c0);
                __syncthreads();   // M's previous readers are done
                sum_stage<EC>(a, s0 + g, e0, ne, c0, nr, (nr + 3) / 4, M, tid);
                __syncthreads();
                sum_mfma_window<NT>(M, c0, (nr + 3) / 4, lane, ra, wa, acc);
            }
        }
        __syncthreads();   // M and red are rewritten by the next group
    }
    // D[4 (l / 16) + i][l % 16] of m-tile m: pixels 4 (l / 16) .. +3 of tile
    // row 4 wv + m, energy 16 n + l % 16
    const int hh = th * kTileH + 4 * wv;
    const int wq = tw * kTileW + 4 * (lane >> 4);
    const bool vec = (a.W & 3) == 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (hh + m >= a.H) break;
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int e = 16 * n + (lane & 15);
            if (e >= ne) continue;
            float* po = a.sum + (size_t)(e0 + e) * plane + (size_t)(hh + m) * a.W + wq;
            if (vec && wq + 3 < a.W) {
                float4 y = *reinterpret_cast<float4*>(po);
                y.x += acc[m][n][0];
                y.y += acc[m][n][1];
                y.z += acc[m][n][2];
                y.w += acc[m][n][3];
                *reinterpret_cast<float4*>(po) = y;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (wq + i < a.W) po[i] += acc[m][n][i];
            }
        }
    }
}

// The wide pixels of a summed model's batch: kWideBlocks blocks per source
// and energy chunk, as in the emit.
template <int EC>
__global__ void __launch_bounds__(kEmitThreads)
wide_sum_kernel(CubArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* H = reinterpret_cast<float*>(smem_raw);
    constexpr int kWaves = kEmitThreads / 64;
    wide_pixels<float, EC, float, true, true>(a, blockIdx.x * kWaves + threadIdx.x / 64,
                                              kWideBlocks * kWaves,
                                              H + (threadIdx.x / 64) * kNDPad, blockIdx.y,
                                              blockIdx.z * EC);
}

// The wide pixels of every source: kWideBlocks blocks per source and energy
// chunk (grid (kWideBlocks, S, chunks)), one wave per pixel (wide_pixels).
template <typename T, int EC, typename OT, bool SCALED>
__global__ void __launch_bounds__(kEmitThreads) wide_emit_kernel(CubArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* H = reinterpret_cast<T*>(smem_raw);
    constexpr int kWaves = kEmitThreads / 64;
    wide_pixels<T, EC, OT, SCALED>(a, blockIdx.x * kWaves + threadIdx.x / 64,
                                   kWideBlocks * kWaves, H + (threadIdx.x / 64) * kNDPad,
                                   blockIdx.y, blockIdx.z * EC);
}

// grid: (emit blocks per source, S, energy chunks).  mode kEmitWithWide: the
// wide pixels by kWideBlocks extra blocks per source of the same launch;
// kEmitFar: the far pixels only (the emit's registers then leave room for
// other work on a CU); kEmitWide: the wide pixels' launch alone.
enum { kEmitWithWide = 0, kEmitFar = 1, kEmitWide = 2 };
template <typename T, int EC, typename OT, bool SCALED>
static int launch_emit(hipStream_t st, const CubArgs& ca, dim3 grid, bool sehw, int mode) {
    if (mode == kEmitWide) {
        const size_t smw = (size_t)(kEmitThreads / 64) * kNDPad * sizeof(T);
        XTS_TRY(allow_lds((const void*)wide_emit_kernel<T, EC, OT, SCALED>, smw));
        wide_emit_kernel<T, EC, OT, SCALED><<<dim3(kWideBlocks, grid.y, grid.z), kEmitThreads,
                                              smw, st>>>(ca);
        XTS_HIP(hipGetLastError());
        return XTS_OK;
    }
    const size_t sm = (size_t)kND * (sehw ? EC + 1 : EC) * sizeof(T);
    if (mode == kEmitFar) {
        const void* fn = sehw ? (const void*)emit_sehw_kernel<T, EC, OT, SCALED, false>
                              : (const void*)emit_shwe_kernel<T, EC, OT, SCALED, false>;
        XTS_TRY(allow_lds(fn, sm));
        if (sehw) emit_sehw_kernel<T, EC, OT, SCALED, false><<<grid, kEmitThreads, sm, st>>>(ca);
        else emit_shwe_kernel<T, EC, OT, SCALED, false><<<grid, kEmitThreads, sm, st>>>(ca);
    } else {
        const dim3 g2(grid.x + kWideBlocks, grid.y, grid.z);
        const void* fn = sehw ? (const void*)emit_sehw_kernel<T, EC, OT, SCALED, true>
                              : (const void*)emit_shwe_kernel<T, EC, OT, SCALED, true>;
        XTS_TRY(allow_lds(fn, sm));
        if (sehw) emit_sehw_kernel<T, EC, OT, SCALED, true><<<g2, kEmitThreads, sm, st>>>(ca);
        else emit_shwe_kernel<T, EC, OT, SCALED, true><<<g2, kEmitThreads, sm, st>>>(ca);
    }
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

// Pooled near pixels (wide = false) or wide pixels (wide = true: a few
// waves per source scan the list).
template <typename T, int EC, typename OT, bool SCALED, bool SUM = false>
static int launch_near_emit(hipStream_t st, const CubArgs& ca, dim3 grid) {
    const size_t sm = (size_t)kND * EC * sizeof(T);
    XTS_TRY(allow_lds((const void*)near_emit_kernel<T, EC, OT, SCALED, SUM>, sm));
    near_emit_kernel<T, EC, OT, SCALED, SUM><<<grid, kNearThreads, sm, st>>>(ca);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

template <typename T, int EC>
static int near_emit(hipStream_t st, const CubArgs& ca, int64_t gx, bool scaled) {
    const dim3 grid((unsigned)gx, (unsigned)ca.S, (unsigned)div_up(ca.E, EC));
    if (ca.out_f32)
        return scaled ? launch_near_emit<T, EC, float, true>(st, ca, grid)
                      : launch_near_emit<T, EC, float, false>(st, ca, grid);
    return scaled ? launch_near_emit<T, EC, double, true>(st, ca, grid)
                  : launch_near_emit<T, EC, double, false>(st, ca, grid);
}

// The map values of the pass-1 pixels and the wide pixels (launch_emit's
// modes).
template <typename T, int EC>
static int emit(hipStream_t st, const CubArgs& ca, dim3 grid, bool scaled,
                int mode = kEmitWithWide) {
    const bool sehw = ca.layout_sehw != 0;
    if (ca.out_f32)
        return scaled ? launch_emit<T, EC, float, true>(st, ca, grid, sehw, mode)
                      : launch_emit<T, EC, float, false>(st, ca, grid, sehw, mode);
    return scaled ? launch_emit<T, EC, double, true>(st, ca, grid, sehw, mode)
                  : launch_emit<T, EC, double, false>(st, ca, grid, sehw, mode);
}

// ------------------------------------------------------------------ factors
// model_map.map_integral.cxx:41-44 (reciprocal), map_correction_factor.cxx:
// 24-50 and scale_map_by_exposure: f[s, e] = exposure x correction.
// map_correction_factor.cxx:24-50 for one (source, energy): the PSF
// integral to the node below the radius plus its whole next segment
// (capped at 1), x 1 / map integral; 1 outside the field of view.
// With exact (XTS_CF_EXACT_RADIUS) the FermiTools-consistent variant: the
// node below the radius for every radius, and the segment integrated to the
// radius itself.
__device__ inline double correction_factor(const SrcGeom& g, double invmi,
                                           const double* __restrict__ upsf,
                                           const double* __restrict__ parint,
                                           const double* __restrict__ sep, int s, int e, int E,
                                           int exact = 0) {
    if (!g.in_fov) return 1.0;
    const double r = g.radius_deg;
    int d;
    if (!exact) {
        d = r < 1.0 ? 1 : (int)floor(1.0 + log(1e4 * r) * kRecipStep);
    } else {
        d = r < 1e-4 ? 0 : (int)floor(1.0 + log(1e4 * r) * kRecipStep);
        while (d + 1 < kND && sep[d + 1] <= r) ++d;   // rounding at a node
        while (d > 0 && sep[d] > r) --d;
    }
    if (d >= kND - 2) d = kND - 2;
    const double t1 = sep[d] * kD2R, t2 = sep[d + 1] * kD2R;
    const double te = exact ? r * kD2R : t2;
    const double y1 = kTwoPi * upsf[((size_t)s * kND + d) * E + e] * sin(t1);
    const double y2 = kTwoPi * upsf[((size_t)s * kND + d + 1) * E + e] * sin(t2);
    const double m = (y2 - y1) / (t2 - t1);
    const double b = y1 - t1 * m;
    const double v = 0.5 * m * (te * te - t1 * t1) + b * (te - t1);
    double integ = v + parint[((size_t)s * kND + d) * E + e];
    integ = integ > 1.0 ? 1.0 : integ;
    return integ * invmi;
}

// map_integral.cxx:21-44 through the row histogram: the in-radius sum of
// the (solid-angle scaled) map is sum_d hist[s, d] uPsf[s, d, e] (FP64
// table), listed pixels included; then the correction factor of each
// (source, energy) from it.  One block per source: each thread takes one
// energy and every P-th row (consecutive threads read consecutive
// energies), the P partial sums are added in LDS.
constexpr int kHistThreads = 1024;
__global__ void __launch_bounds__(kHistThreads)
hist_integral_kernel(int E, const double* __restrict__ hist, const double* __restrict__ upsf,
                     double* __restrict__ mapint, const SrcGeom* __restrict__ geo,
                     const double* __restrict__ parint, const double* __restrict__ sep,
                     double* __restrict__ cf_out, int exact) {
    __shared__ double h_s[kNDPad];
    __shared__ double part[kHistThreads];
    const int s = blockIdx.x, tid = threadIdx.x;
    for (int d = tid; d < kND; d += kHistThreads) h_s[d] = hist[(size_t)s * kND + d];
    const int Ec = min(E, kHistThreads);
    const int P = kHistThreads / Ec;
    const int el = tid % Ec, ph = tid / Ec;
    const double* u = upsf + (size_t)s * kND * E;
    const SrcGeom g = geo[s];
    for (int eb = 0; eb < E; eb += Ec) {
        __syncthreads();
        const int e = eb + el;
        double acc = 0.0;
        if (ph < P && e < E) {
#pragma unroll 4
            for (int d = ph; d < kND; d += P) acc = fma(h_s[d], u[(size_t)d * E + e], acc);
        }
        part[tid] = acc;
        __syncthreads();
        if (ph == 0 && e < E) {
            double tot = 0.0;
            for (int q = 0; q < P; ++q) tot += part[q * Ec + el];
            mapint[(size_t)s * E + e] = tot;
            const double invmi = tot != 0.0 ? 1.0 / tot : tot;
            cf_out[(size_t)s * E + e] =
                correction_factor(g, invmi, upsf, parint, sep, s, e, E, exact);
        }
    }
}

__global__ void factor_kernel(int S, int E, const SrcGeom* __restrict__ geo,
                              const double* __restrict__ mapint,
                              const double* __restrict__ invmi_in, const double* __restrict__ upsf,
                              const double* __restrict__ parint, const double* __restrict__ sep,
                              const double* __restrict__ expo, double* __restrict__ cf_out,
                              double* __restrict__ invmi_out, int s_begin = 0) {
    const int k = s_begin * E + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S * E) return;
    const int s = k / E, e = k % E;
    const SrcGeom g = geo[s];
    double invmi = 0.0;
    if (mapint) {
        const double m = mapint[k];
        invmi = m != 0.0 ? 1.0 / m : m;
    } else if (invmi_in) {
        invmi = invmi_in[k];
    }
    const double cf = correction_factor(g, invmi, upsf, parint, sep, s, e, E);
    if (cf_out) cf_out[k] = cf;
    if (invmi_out) invmi_out[k] = invmi;
}

// pixel-factor scaling on a double [S,H,W,E] map (solid angle)
__global__ void scale_pixel_kernel(double* __restrict__ map, size_t SHW, int HW, int E,
                                   const double* __restrict__ f) {
    const size_t n = SHW * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        map[k] *= f[(k / E) % HW];
}

__global__ void scale_se_kernel(double* __restrict__ map, size_t HW, int S, int E,
                                const double* __restrict__ f) {
    const size_t n = (size_t)S * HW * E;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t s = k / (HW * E);
        map[k] *= f[s * E + (k % E)];
    }
}

// map_integral.cxx:21-44 on a double [S,H,W,E] map: one block per source
__global__ void map_integral_kernel(int H, int W, int E, const double* __restrict__ map,
                                    const double* __restrict__ centres,
                                    const SrcGeom* __restrict__ geo, double* __restrict__ out) {
    const int s = blockIdx.x;
    const SrcGeom g = geo[s];
    for (int e = threadIdx.x; e < E; e += blockDim.x) out[(size_t)s * E + e] = 0.0;
    if (!g.in_fov) return;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        double acc = 0.0;
        for (int h = 0; h < H; ++h)
            for (int w = 0; w < W; ++w) {
                const double* c = centres + 3 * (h * (W + 1) + w);
                if (dir_diff(g.dir, c) <= kD2R * g.radius_deg)
                    acc += map[(((size_t)s * H + h) * W + w) * E + e];
            }
        out[(size_t)s * E + e] = acc != 0.0 ? 1.0 / acc : acc;
    }
}

__global__ void pix2sph_kernel(Wcs w, int64_t n, const double* __restrict__ in,
                               double* __restrict__ out, int inverse) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (inverse) wcs_sph2pix(w, in[2 * k], in[2 * k + 1], out + 2 * k, out + 2 * k + 1);
        else wcs_pix2sph(w, in[2 * k], in[2 * k + 1], out + 2 * k, out + 2 * k + 1);
    }
}

// ------------------------------------------------------------------ host side
struct ImageGeom {
    double* corners = nullptr;
    double* centres = nullptr;
    double* omega = nullptr;
    SrcGeom* geo = nullptr;
};

static int check_wcs(const xts_wcs* wcs) {
    XTS_CHECK(wcs != nullptr, "null wcs");
    XTS_CHECK(wcs->naxis1 > 0 && wcs->naxis2 > 0 && wcs->naxis1 <= 65536 && wcs->naxis2 <= 65536,
              "bad image size %lld x %lld", (long long)wcs->naxis1, (long long)wcs->naxis2);
    XTS_CHECK(wcs->proj == XTS_PROJ_CAR || wcs->proj == XTS_PROJ_AIT, "unknown projection %d",
              wcs->proj);
    XTS_CHECK(wcs->cdelt1 != 0.0 && wcs->cdelt2 != 0.0, "zero CDELT");
    return XTS_OK;
}

// Geometry + (optionally) per-source radius into the kSlotGeom slot.
static int build_geometry(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph_dev, int S,
                          ImageGeom* ig, bool with_radius = true, bool keep = false) {
    const int H = (int)wcs->naxis1, W = (int)wcs->naxis2;
    const size_t ncorn = (size_t)(H + 1) * (W + 1);
    const size_t bytes = ncorn * 3 * 8 * 2 + (size_t)H * W * 8;
    void* base;
    XTS_TRY(slot_reserve(ctx, kSlotGeom, bytes, &base));
    char* p = (char*)base;
    ig->corners = (double*)p; p += ncorn * 3 * 8;
    ig->centres = (double*)p; p += ncorn * 3 * 8;
    ig->omega = (double*)p;
    void* g;
    XTS_TRY(slot_reserve(ctx, kSlotSrcGeo, (size_t)std::max(S, 1) * sizeof(SrcGeom), &g));
    ig->geo = (SrcGeom*)g;
    const Wcs w = wcs_setup(*wcs);
    // the pixel geometry depends on the WCS alone: recomputed only when the
    // image changes (the slot keeps the last image's); always inside an
    // arena, whose contents do not outlive the call
    // keep: the slot holds this image's geometry from the call just before
    // (the batched engine's emit half)
    const bool cached = ctx->arena.base == nullptr && ctx->geom_cache;
    if (!keep &&
        (!cached || !ctx->geom_valid || std::memcmp(&ctx->geom_wcs, wcs, sizeof(xts_wcs)) != 0)) {
        geometry_kernel<<<std::min<unsigned>(div_up(ncorn, 256), 4096), 256, 0, ctx->stream>>>(
            w, H, W, ig->corners, ig->centres);
        solid_angle_kernel<<<std::min<unsigned>(div_up((int64_t)H * W, 256), 4096), 256, 0,
                             ctx->stream>>>(H, W, ig->centres, ig->omega);
        if (cached) {
            ctx->geom_wcs = *wcs;
            ctx->geom_valid = true;
        }
    }
    if (S > 0 && src_sph_dev && with_radius)
        radius_kernel<<<S + 1, 256, 0, ctx->stream>>>(w, H, W, S, src_sph_dev, ig->corners, ig->geo);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

int image_solid_angle(xts_ctx* ctx, const xts_wcs* wcs, const double** omega) {
    XTS_TRY(check_wcs(wcs));
    // the geometry slots are the pending first half's of a split model map
    ctx->pending_rows.valid = false;
    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, nullptr, 0, &ig));
    *omega = ig.omega;
    return XTS_OK;
}

// The PSF table in the cubature's precision, energies padded to Ep (a
// multiple of 4); the same threads clear the nzero accumulators of the
// model map.
__device__ inline void lut_convert_part(const double* __restrict__ upsf, int S, int E, int Ep,
                                        float* __restrict__ out32, double* __restrict__ out64,
                                        double* __restrict__ zero, size_t nzero, size_t i0,
                                        size_t step) {
    const size_t n = (out32 || out64) ? (size_t)S * kND * Ep : 0;
    for (size_t k = i0; k < n; k += step) {
        const int e = (int)(k % Ep);
        const size_t sd = k / Ep;
        const double v = e < E ? upsf[sd * E + e] : 0.0;
        if (out32) out32[k] = (float)v;
        if (out64) out64[k] = v;
    }
    for (size_t k = i0; k < nzero; k += step) zero[k] = 0.0;
}

// Both in one launch (they are independent): blocks 0 .. S take the
// boundary radii and field-of-view flags, the rest the table conversion.
__global__ void __launch_bounds__(256)
radius_lut_kernel(Wcs w, int H, int W, int S, int nrad, const double* __restrict__ src_sph,
                  const double* __restrict__ corners, SrcGeom* __restrict__ geo,
                  const double* __restrict__ upsf, int E, int Ep, float* __restrict__ out32,
                  double* __restrict__ out64, double* __restrict__ zero, size_t nzero) {
    // blocks 0 .. nrad-1 (nrad = S + 1 or 0): radii and field-of-view flags
    if ((int)blockIdx.x < nrad) {
        radius_block(w, H, W, S, src_sph, corners, geo, blockIdx.x);
        return;
    }
    lut_convert_part(upsf, S, E, Ep, out32, out64, zero, nzero,
                     (size_t)(blockIdx.x - nrad) * blockDim.x + threadIdx.x,
                     (size_t)(gridDim.x - nrad) * blockDim.x);
}

// The rule as per-point records on the device, with its bilinear-weight
// bound, weight sum and grid structure; re-uploaded only when its content
// changes.
static int upload_rule(xts_ctx* ctx, const double* rule, int64_t N, RuleUVW** dev,
                       const float** lean, double* rho_scale, double* wsum, int* tnu, int* tnv) {
    const double* r = rule ? rule : kSquareDeg7;
    if (!rule) N = 12;
    XTS_CHECK(N > 0 && N <= 1024, "rule size %lld out of range (1..1024)", (long long)N);
    // inside an arena the rule is uploaded for this call only: the context's
    // cache is neither used nor changed
    const bool arena = ctx->arena.base != nullptr;
    xts_rule_cache in_arena;
    xts_rule_cache& rc = arena ? in_arena : ctx->rule;
    if (arena || !(rc.dev && rc.host.size() == (size_t)(3 * N) &&
                   std::memcmp(rc.host.data(), r, (size_t)(3 * N) * sizeof(double)) == 0)) {
        std::vector<RuleUVW> rp((size_t)N);
        double rs = 1.0;
        for (int64_t p = 0; p < N; ++p) {
            const double u = r[3 * p], v = r[3 * p + 1];
            rp[p] = {u, v, 0.25 * r[3 * p + 2]};
            // sum of |bilinear corner weights|: 1 inside the square
            const double b = 0.25 * (std::fabs((1 - u) * (1 - v)) + std::fabs((1 + u) * (1 - v)) +
                                     std::fabs((1 + u) * (1 + v)) + std::fabs((1 - u) * (1 + v)));
            rs = std::max(rs, b);
        }
        double ws = 0.0;
        for (auto& q : rp) ws += q.wq;
        // grid rule in row-major order: u depends on p / nv only, v on p % nv only
        int nu = 0, nv = 0;
        for (int64_t p = 1; p < N && nv == 0; ++p)
            if (rp[p].u != rp[0].u) nv = (int)p;
        if (nv > 0 && N % nv == 0 && nv % 4 == 0) {
            nu = (int)(N / nv);
            bool grid = true;
            for (int64_t p = 0; p < N && grid; ++p)
                grid = rp[p].u == rp[(p / nv) * nv].u && rp[p].v == rp[p % nv].v;
            if (!grid) nu = nv = 0;
        } else {
            nv = 0;
        }
        // the records, then the FP32 arrays of the lean point loop, formed
        // in FP32 as the kernels use them, zero past the last point
        const int64_t N4 = (N + 3) / 4 * 4;
        const size_t rec = ((size_t)N * sizeof(RuleUVW) + 255) & ~size_t(255);
        std::vector<unsigned char> blob(rec + (size_t)kLeanArrays * N4 * sizeof(float), 0);
        std::memcpy(blob.data(), rp.data(), (size_t)N * sizeof(RuleUVW));
        float* la = reinterpret_cast<float*>(blob.data() + rec);
        for (int64_t p = 0; p < N; ++p) {
            const float wq = (float)rp[p].wq;
            la[p] = (float)rp[p].u;
            la[N4 + p] = (float)rp[p].v;
            la[2 * N4 + p] = wq;
        }
        void* d;
        if (arena) {
            ctx->arena_rule_blob = std::move(blob);
            XTS_TRY(to_device(ctx, kSlotRule, ctx->arena_rule_blob.data(),
                              ctx->arena_rule_blob.size(), &d));
        } else {
            XTS_TRY(to_device(ctx, kSlotRule, blob.data(), blob.size(), &d));
        }
        rc.host.assign(r, r + 3 * N);
        rc.dev = d;
        ++ctx->generation;
        rc.lean_off = rec;
        rc.rho_scale = rs;
        rc.wsum = ws;
        rc.nu = nu;
        rc.nv = nv;
    }
    *dev = (RuleUVW*)rc.dev;
    *lean = reinterpret_cast<const float*>((const char*)rc.dev + rc.lean_off);
    *rho_scale = rc.rho_scale;
    *wsum = rc.wsum;
    *tnu = rc.nu;
    *tnv = rc.nv;
    return XTS_OK;
}

// The device pipeline shared by the host and _dev entry points.
//   full = true : model map (x omega, x exposure x correction)
//   full = false: raw cubature (cubature_convolve_map_psf)
// Phases of one model map: kRows needs only the image, the sources and the
// rule (geometry, radii, row weights, near-pixel rows); kFinish the PSF
// table, exposure and output (table conversion, map integral and factor,
// emit).  kAll runs both in one call.  The batched engine splits the second
// half once more: kPhaseRows | kPhaseTable, then kPhaseEmit alone (the
// geometry and table of the first call kept).
enum { kPhaseRows = 1, kPhaseTable = 2, kPhaseEmit = 4, kPhaseFinish = 6, kPhaseAll = 7,
       kPhaseSum = 8 };   // in place of kPhaseEmit: the batch added x flux to a summed model

// Per-source int counters (near and wide pixels), in doubles of the
// accumulator slot.
static size_t model_map_counter_doubles(int64_t S) { return (size_t)S; }

// Source batches of the pipelined map for this call, or 1 (the passes over
// all sources at once): the FP32 cubature into an FP32 [S, E, H, W] map,
// full model map, not from a caller's arena; tuning key
// "pipeline_batches" (0: the default, 1: off, n: n batches).  By default a
// map of at least kPipeMinValues values runs in kPipeDefaultBatches
// batches (2,000 sources x 37 x 400 x 400: 15.1 ms against 16.4 for the
// same kernels one after the other, on one MI355X), a surrogate map in
// kPipeSurBatches (its first batch's row weights and network, and its last
// emit, run with nothing beside them: 12 batches took 17.9-18.75 ms against
// 18.1-18.97 for 8 in four same-box comparisons; 19.1-20.0 as one pass).
constexpr int64_t kPipeDefaultBatches = 8;
constexpr int64_t kPipeSurBatches = 12;
constexpr int64_t kPipeMinSources = 16;   // sources per batch at least
constexpr double kPipeMinValues = 4e9;
static int64_t pipeline_batches(const xts_ctx* ctx, int64_t S, int64_t plane, int64_t E,
                                int32_t flags, bool full, int phase, int sur) {
    const bool ok = full && !(flags & XTS_COMPUTE_F64) && (flags & XTS_OUT_F32) &&
                    (flags & XTS_LAYOUT_SEHW) && !(phase & kPhaseSum) &&
                    ctx->arena.base == nullptr && phase != kPhaseEmit;
    if (!ok) return 1;
    int64_t B = ctx->tune_pipe_batches;
    if (B == 0)
        B = (double)S * plane * E >= kPipeMinValues ? (sur ? kPipeSurBatches : kPipeDefaultBatches)
                                                    : 1;
    B = std::min<int64_t>(B, S / kPipeMinSources);
    return B > 1 ? B : 1;
}

// Sources per block of the tile-frame row-weight pass: at most 64, fewer
// when that leaves under about eight blocks per CU.
static int64_t rw_frames_per_block(const xts_ctx* ctx, int64_t S, int64_t ntiles) {
    if (ctx->tune_rw_frames > 0) return ctx->tune_rw_frames;
    const int64_t fill = std::max<int64_t>(1, S * ntiles / ((int64_t)ctx->cu_count * 8));
    return std::max<int64_t>(4, std::min<int64_t>(64, fill));
}

// Launches the tile-frame row-weight pass over the sources of c on stream
// st: items of one tile and a run of about `per_block` sources, one block
// per item, or per_cu resident blocks per CU striding over the items.
// Launches the surrogate's network pass over the sources of c on stream st:
// blocks striding over each source's tiles (about 4 x want_blocks in all),
// or per_cu resident blocks per CU striding over (source, run of four tiles)
// items.
static int launch_sur(xts_ctx* ctx, const CubArgs& c, int64_t want_blocks, int64_t per_cu,
                      hipStream_t st) {
    const int64_t tiles = (int64_t)c.tiles_w * c.tiles_h;
    const int64_t runs = div_up(tiles, kSurThreads / 64);
    dim3 grid((unsigned)std::min<int64_t>(std::max<int64_t>(1, 4 * want_blocks / c.S), runs),
              (unsigned)c.S);
    if (per_cu > 0) grid = dim3((unsigned)std::min<int64_t>(c.S * runs, per_cu * ctx->cu_count));
    const int resident = per_cu > 0 ? 1 : 0;
    if (c.sur == 3) sur_rows_kernel<MlpPacked8><<<grid, kSurThreads, 0, st>>>(c, resident);
    else if (ctx->tune_sur_lo) sur_rows_kernel<MlpPacked><<<grid, kSurThreads, 0, st>>>(c, resident);
    else sur_rows_kernel<MlpPackedH><<<grid, kSurThreads, 0, st>>>(c, resident);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

static int launch_rw_frames(xts_ctx* ctx, const CubArgs& c, int64_t per_block, int64_t per_cu,
                            hipStream_t st) {
    const int64_t ntiles = (int64_t)c.tiles_w * c.tiles_h;
    const int64_t nsplit = std::max<int64_t>(1, (c.S + per_block - 1) / per_block);
    const int64_t nitem = ntiles * nsplit;
    const int64_t grid = per_cu > 0 ? std::min<int64_t>(nitem, per_cu * ctx->cu_count) : nitem;
    if (c.sur) rowweight_frames_kernel<true><<<(unsigned)grid, kThreads, 0, st>>>(c, (int)nsplit);
    else rowweight_frames_kernel<false><<<(unsigned)grid, kThreads, 0, st>>>(c, (int)nsplit);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

// The pipelined map (FP32 [S, E, H, W], B source batches of Sb): a second
// stream computes the row weights and near rows of batch b + 1 (with the
// surrogate, and its network pass: tuning surrogate_beside) while the
// context stream takes batch b's correction factors, map values and near
// pixels.  The row-weight pass (rowweight_frames_kernel) reads no global
// memory per source, so it keeps its pace beside the store-bound emit.  Row
// weights and lists of two batches alternate between two buffers (parity
// b & 1): batch b + 1's row weights wait for batch b - 1's map.  kPhaseRows
// enqueues batch 0's row weights (the radii and cleared accumulators are
// already enqueued on the context stream); the emit phase the rest.  ca
// describes all S sources, with the two batches' row buffers at its list
// pointers.
static int run_pipeline(xts_ctx* ctx, const CubArgs& ca, int64_t B, int64_t Sb, int phase,
                        const double* upsf, const double* parint, const double* sep, int exact,
                        size_t pool_cap) {
    const int64_t S = ca.S, E = ca.E;
    const size_t plane = (size_t)ca.H * ca.W;
    const int64_t want_blocks = (int64_t)ctx->cu_count * 4;
    if (!ctx->pipe_stream)
        XTS_HIP(hipStreamCreateWithFlags(&ctx->pipe_stream, hipStreamNonBlocking));
    while ((int64_t)ctx->pipe_ev.size() < 2 * B + 1) {
        hipEvent_t e;
        XTS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->pipe_ev.push_back(e);
    }
    // evR[b]: batch b's row weights and near rows done (R); evM[b]: batch b
    // written (M)
    hipEvent_t* evR = ctx->pipe_ev.data();
    hipEvent_t* evM = evR + B;
    hipEvent_t fork = evR[2 * B];
    const hipStream_t M = ctx->stream, R = ctx->pipe_stream;
    const int64_t per_block =
        rw_frames_per_block(ctx, std::min(Sb, S), (int64_t)ca.tiles_w * ca.tiles_h);
    // resident row-weight blocks per CU beside the emit: two leave room for
    // one emit block (LDS 2 x 31 + 66 KB); more crowd the emit out
    const int64_t per_cu = ctx->tune_rw_frames_cu > 0 ? ctx->tune_rw_frames_cu : 2;
    auto batch = [&](int64_t b) {
        CubArgs c = ca;
        const int64_t b0 = b * Sb;
        const size_t par = (size_t)(b & 1) * (size_t)Sb * plane;
        c.S = (int)std::min(Sb, S - b0);
        c.tiles_per_block = 1;
        c.rw_persist = 0;
        c.geo = ca.geo + b0;
        c.lut = (const float*)ca.lut + (size_t)b0 * kND * ca.Ep;
        c.near_count = ca.near_count + b0;
        c.wide_count = ca.wide_count + b0;
        c.hist = ca.hist + (size_t)b0 * kND;
        c.mapint = ca.mapint + (size_t)b0 * E;
        c.cf = ca.cf + (size_t)b0 * E;
        c.expo = ca.expo ? ca.expo + (size_t)b0 * E : nullptr;
        c.out = ca.out ? (void*)((float*)ca.out + (size_t)b0 * E * plane) : nullptr;
        c.rw = (float4*)ca.rw + par;
        c.rw_row = ca.rw_row + par;
        c.near_list = ca.near_list + par;
        c.near_off = ca.near_off + par;
        c.near_rr = ca.near_rr + par;
        c.wide_list = ca.wide_list + par;
        c.near_pool = (float*)ca.near_pool + (size_t)(b & 1) * (size_t)Sb * pool_cap;
        if (ca.sur_list) {
            const size_t st = (size_t)(b & 1) * (size_t)Sb * (size_t)ca.tiles_w * ca.tiles_h;
            c.sur_list = ca.sur_list + 2 * st * kThreads;
            c.sur_count = ca.sur_count + st;
        }
        c.emit_gx = (int)div_up(plane, kEmitBlockPix);
        c.emit_prio = (int)ctx->tune_emit_prio;
        return c;
    };
    const size_t tsz = 4;
    const int N4 = (ca.N + 3) & ~3;
    const size_t sm_near = (size_t)kNDPad * tsz + (size_t)N4 * 4 * tsz + (size_t)N4 * 16 +
                           (size_t)kNearRows * kNearThreads * tsz;
    XTS_TRY(allow_lds((const void*)near_rows_kernel<float>, sm_near));
    auto near_gx = [&](const CubArgs& c) {
        return std::min<int64_t>(std::max<int64_t>(1, 2 * want_blocks / c.S),
                                 div_up(plane, kNearPix));
    };
    // the near pixels' map values: blocks per source, each staging the
    // source's table once (2: 0.38 ms per map at 2,000 sources on 400x400,
    // against 0.62 for a grid as wide as the near rows')
    auto near_emit_gx = [&](const CubArgs& c) {
        const int64_t want = ctx->tune_near_emit_gx > 0 ? ctx->tune_near_emit_gx : 2;
        return std::min<int64_t>(want, near_gx(c));
    };
    // batch b's row weights and near rows on R
    auto rows = [&](int64_t b) -> int {
        const CubArgs c = batch(b);
        // the first batch's row weights run alone: one block per item
        XTS_TRY(launch_rw_frames(ctx, c, per_block, b == 0 ? 0 : per_cu, R));
        // the network pass on R as surrogate_beside resident blocks per CU
        // (4 by default: with 3 to 6 the map takes 19.0-19.1 ms at 2,000
        // sources x 37 x 400 x 400, with 2 20.9, with 1 30.9; on M, 20.0;
        // the network needs the occupancy, so the gain is the surrogate row
        // weights beside the emit)
        if (c.sur && ctx->tune_sur_beside > 0)
            XTS_TRY(launch_sur(ctx, c, want_blocks, b == 0 ? 0 : ctx->tune_sur_beside, R));
        near_rows_kernel<float><<<dim3((unsigned)near_gx(c), (unsigned)c.S), kNearThreads,
                                  sm_near, R>>>(c);
        XTS_HIP(hipGetLastError());
        XTS_HIP(hipEventRecord(evR[b], R));
        return XTS_OK;
    };
    // batch b's factors, map values and near pixels on M
    auto emit_batch = [&](int64_t b) -> int {
        const CubArgs c = batch(b);
        const size_t b0 = (size_t)(b * Sb);
        XTS_HIP(hipStreamWaitEvent(M, evR[b], 0));
        if (c.sur && ctx->tune_sur_beside == 0) {
            // the network pass at full occupancy on M
            TimedRegion tr(ctx, "surrogate");
            XTS_TRY(launch_sur(ctx, c, want_blocks, 0, M));
        }
        {
            TimedRegion tr(ctx, "factor");
            hist_integral_kernel<<<(unsigned)c.S, kHistThreads, 0, M>>>(
                (int)E, c.hist, upsf + b0 * kND * E, c.mapint, c.geo, parint + b0 * kND * E,
                sep, const_cast<double*>(c.cf), exact);
            XTS_HIP(hipGetLastError());
        }
        // energies per emit block: 40 (66 KB of LDS for the table) or 20
        // (34 KB: two emit blocks fit beside the row-weight blocks)
        // the wide pixels as extra emit blocks (124 VGPRs: two row-weight
        // waves of 96 and two emit waves still fit a SIMD), or (tuning
        // emit_wide_apart) a launch of their own after a far-pixel emit of 70
        // VGPRs: the same time per step
        const bool apart = ctx->tune_emit_wide_apart != 0;
        for (const int mode : {apart ? kEmitFar : kEmitWithWide, kEmitWide}) {
            if (mode == kEmitWide && !apart) break;
            TimedRegion tr(ctx, mode == kEmitWide ? "wide" : "emit");
            if (ctx->tune_emit_ec == 20) {
                const dim3 grid((unsigned)c.emit_gx, (unsigned)c.S, (unsigned)div_up(E, 20));
                XTS_TRY((emit<float, 20>(M, c, grid, true, mode)));
            } else {
                const dim3 grid((unsigned)c.emit_gx, (unsigned)c.S,
                                (unsigned)div_up(E, kEmitEC32));
                XTS_TRY((emit<float, kEmitEC32>(M, c, grid, true, mode)));
            }
        }
        {
            TimedRegion tr(ctx, "near_emit");
            XTS_TRY((near_emit<float, 40>(M, c, near_emit_gx(c), true)));
        }
        XTS_HIP(hipEventRecord(evM[b], M));
        return XTS_OK;
    };
    if (phase & kPhaseRows) {
        // R starts after the radii and cleared accumulators on M
        XTS_HIP(hipEventRecord(fork, M));
        XTS_HIP(hipStreamWaitEvent(R, fork, 0));
        XTS_TRY(rows(0));
    }
    if (!(phase & kPhaseEmit)) return XTS_OK;
    for (int64_t b = 0; b < B; ++b) {
        XTS_TRY(emit_batch(b));
        if (b + 1 < B) {
            // batch b + 1 reuses the buffers of batch b - 1; R runs up to one
            // batch ahead of M (waiting for batch b itself, as measured,
            // leaves M idle between batches)
            if (b >= 1) XTS_HIP(hipStreamWaitEvent(R, evM[b - 1], 0));
            XTS_TRY(rows(b + 1));
        }
    }
    // M's last wait was on evR[B - 1]: R's work is joined into M
    return XTS_OK;
}

// pipe: the map may run as the pipelined source batches (run_pipeline);
// false for callers that split a map into phases of their own (the batched
// engine).
static int model_map_device(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph, int64_t S,
                            const double* expo, const double* upsf, const double* parint,
                            int64_t E, const double* rule, int64_t N, int32_t flags, void* out,
                            bool full, int phase = kPhaseAll, const double* flux = nullptr,
                            bool pipe = true) {
    // any model map other than the second half of a split one overwrites the
    // near lists, counters and radii that a pending first half left: its
    // second half is refused from now on
    if (phase != kPhaseFinish) ctx->pending_rows.valid = false;
    XTS_TRY(check_wcs(wcs));
    // (S E < 2^31: source-energy indices are 32-bit in the kernels)
    XTS_CHECK(S > 0 && S <= (1 << 24) && E > 0 && E <= 4096 && S * E < (1LL << 31),
              "bad S=%lld or E=%lld", (long long)S, (long long)E);
    XTS_CHECK(src_sph, "null device pointer");
    XTS_CHECK(!(phase & kPhaseTable) || upsf, "null device pointer");
    XTS_CHECK(!(phase & (kPhaseEmit | kPhaseSum)) || out, "null device pointer");
    XTS_CHECK(!(phase & kPhaseSum) || (flux && !(phase & kPhaseEmit) && full &&
                                       !(flags & XTS_COMPUTE_F64) && (flags & XTS_OUT_F32)),
              "a summed model needs the flux, the FP32 cubature and no per-source output");
    XTS_CHECK(!(phase & kPhaseTable) || !full || (expo && parint), "null exposure/parint");
    XTS_CHECK((flags & ~0x7F) == 0, "unknown flags 0x%x", flags);
    const int H = (int)wcs->naxis1, W = (int)wcs->naxis2;
    XTS_CHECK((int64_t)H * W <= (int64_t)1 << 30, "image of %lld pixels exceeds 2^30",
              (long long)H * W);
    const bool f64 = (flags & XTS_COMPUTE_F64) != 0;
    // 1 the BF16 network, 2 no network, 3 the FP8 network
    const int sur = (flags & XTS_SURROGATE)
                        ? ((flags & XTS_SURROGATE_NO_NET) ? 2 : (flags & XTS_SURROGATE_FP8) ? 3 : 1)
                        : 0;
    XTS_CHECK(!(flags & XTS_SURROGATE_FP8) || (sur && !(flags & XTS_SURROGATE_NO_NET)),
              "XTS_SURROGATE_FP8 needs XTS_SURROGATE (and the network)");
    XTS_CHECK(!(flags & XTS_SURROGATE_NO_NET) || sur, "XTS_SURROGATE_NO_NET needs XTS_SURROGATE");
    XTS_CHECK(!sur || (!f64 && full), "the surrogate serves the FP32 model map only");
    XTS_CHECK((sur != 1 && sur != 3) || ctx->surrogate_ready,
              "XTS_SURROGATE: no network set (xts_ctx_set_surrogate)");
    const int sehw = (flags & XTS_LAYOUT_SEHW) ? 1 : 0;
    const int of32 = (flags & XTS_OUT_F32) ? 1 : 0;

    // the pipelined map (run_pipeline): the FP32 [S, E, H, W] model map in
    // pipe_b source batches of pipe_sb, row weights of two batches resident
    const int64_t pipe_b = pipe ? pipeline_batches(ctx, S, (int64_t)wcs->naxis1 * wcs->naxis2, E,
                                                   flags, full, phase, sur)
                                : 1;
    const int64_t pipe_sb = pipe_b > 1 ? (S + pipe_b - 1) / pipe_b : S;
    // the radii are launched with the table conversion below; a pipelined
    // second half keeps the first half's geometry (its row weights may still
    // be reading it on the second stream)
    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, src_sph, (int)S, &ig, false,
                           phase == kPhaseEmit || (pipe_b > 1 && !(phase & kPhaseRows))));
    const xts_const_tables* ct;
    XTS_TRY(const_tables(ctx, &ct));
    RuleUVW* rp;
    const float* lean;
    double rho_scale, wsum;
    int tnu, tnv;
    XTS_TRY(upload_rule(ctx, rule, N, &rp, &lean, &rho_scale, &wsum, &tnu, &tnv));
    if (!rule) N = 12;

    // PSF table in the cubature precision
    const int Ep = (int)((E + 3) / 4 * 4);
    void* lut;
    const size_t lut_elems = (size_t)S * kND * Ep;
    XTS_TRY(slot_reserve(ctx, kSlotLut, lut_elems * (f64 ? 8 : 4), &lut));
    // counters [2 S] (int, padded to doubles: near and wide pixels per
    // source) and FP64 accumulators: hist [S][401], mapint [S][E], factor
    // [S][E]; the counters and hist start at zero
    const size_t cnt_d = model_map_counter_doubles(S);
    const size_t nacc = cnt_d + (full ? (size_t)S * kND + 2 * (size_t)S * E : 0);
    void* accp;
    XTS_TRY(slot_reserve(ctx, kSlotAcc, nacc * 8, &accp));
    double* acc = (double*)accp;
    const size_t nzero = full ? cnt_d + (size_t)S * kND : cnt_d;
    if (phase != kPhaseEmit) {
        // radii and cleared accumulators before the row weights; the table
        // once the PSF is known (one launch when both are)
        const bool rad = phase & kPhaseRows, conv = phase & kPhaseTable;
        const unsigned nrad = rad ? (unsigned)(S + 1) : 0u;
        const unsigned nconv = std::min<unsigned>(div_up(conv ? lut_elems : nzero, 256), 8192);
        radius_lut_kernel<<<nrad + std::max(nconv, 1u), 256, 0, ctx->stream>>>(
            wcs_setup(*wcs), H, W, (int)S, (int)nrad, src_sph, ig.corners, ig.geo,
            conv ? upsf : nullptr, (int)E, Ep, conv && !f64 ? (float*)lut : nullptr,
            conv && f64 ? (double*)lut : nullptr, acc, rad ? nzero : 0);
        XTS_HIP(hipGetLastError());
    }

    // sources whose row weights and lists are resident at once
    const int64_t nbuf = pipe_b > 1 ? 2 * pipe_sb : S;
    // per-pixel row weights and first rows; listed (near) pixels, their pool
    // offsets and row ranges
    const size_t plane = (size_t)H * W;
    const size_t rw_bytes = (size_t)nbuf * plane * (f64 ? 32 : 16);
    void* rows;
    XTS_TRY(slot_reserve(ctx, kSlotRows, rw_bytes + (size_t)nbuf * plane * 20 + 512, &rows));
    // pool of listed pixels' row weights, one segment per source with a slot
    // of kNearRows entries for each of the first 4096 list positions; a
    // pixel beyond is evaluated directly (as a wide pixel)
    const int64_t pool_cap = std::min<int64_t>(4096 * kNearRows, ((int64_t)1 << 30) / nbuf);
    void* pool;
    XTS_TRY(slot_reserve(ctx, kSlotNearPool, (size_t)(pool_cap * nbuf) * (f64 ? 8 : 4), &pool));

    CubArgs ca{};
    ca.corners = ig.corners;
    ca.centres = ig.centres;
    ca.omega = full ? ig.omega : nullptr;
    ca.geo = ig.geo;
    ca.lut = lut;
    ca.uvw = rp;
    ca.lean = lean;
    ca.inv_sep = ct->inv32;
    ca.inv_sep64 = ct->inv64;
    ca.rw = rows;
    ca.rw_row = (int*)((char*)rows + rw_bytes);
    ca.near_list = ca.rw_row + (size_t)nbuf * plane;
    ca.near_off = ca.near_list + (size_t)nbuf * plane;
    ca.near_rr = ca.near_off + (size_t)nbuf * plane;
    ca.wide_list = ca.near_rr + (size_t)nbuf * plane;
    ca.near_count = (int*)acc;
    ca.wide_count = ca.near_count + S;
    ca.sur = sur;
    ca.sur_net = sur == 1 ? (const MlpPacked*)ctx->surrogate : nullptr;
    ca.sur_net8 = sur == 3 ? (const MlpPacked8*)ctx->surrogate8 : nullptr;
    ca.sur_neth = sur == 1 ? (const MlpPackedH*)ctx->surrogate_h : nullptr;
    ca.sur_list = nullptr;
    ca.sur_count = nullptr;
    if (sur) {
        // per tile: 256 list slots of 32 bytes and the tile's count
        const size_t nt =
            (size_t)nbuf * ((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH);
        void* sl;
        XTS_TRY(slot_reserve(ctx, kSlotSurList, nt * (kThreads * 32 + 4), &sl));
        ca.sur_list = (float4*)sl;
        ca.sur_count = (int*)((char*)sl + nt * kThreads * 32);
    }
    ca.near_pool = pool;
    ca.near_pool_cap = (int)pool_cap;
    ca.hist = full ? acc + cnt_d : nullptr;
    ca.mapint = full ? ca.hist + (size_t)S * kND : nullptr;
    double* fac = full ? ca.mapint + (size_t)S * E : nullptr;
    ca.expo = full ? expo : nullptr;
    ca.cf = fac;
    ca.out = out;
    ca.S = (int)S; ca.H = H; ca.W = W; ca.E = (int)E; ca.Ep = Ep; ca.N = (int)N;
    ca.tiles_w = (W + kTileW - 1) / kTileW;
    ca.tiles_h = (H + kTileH - 1) / kTileH;
    ca.layout_sehw = sehw;
    ca.out_f32 = of32;
    ca.rho_scale = rho_scale;
    ca.wsum = wsum;
    ca.tensor_nu = tnu;
    ca.tensor_nv = tnv;
    // j = 1 + recipstep (ln 1e4 + ln(2 R2D) + ln 2 log2(asin(chord / 2)))
    ca.j0 = 1.0 + kRecipStep * (kLn1e4 + std::log(2.0 * kR2D));
    ca.jk = kRecipStep * 0.69314718055994530942;

    if (pipe_b > 1)
        return run_pipeline(ctx, ca, pipe_b, pipe_sb, phase, upsf, parint, ct->sep,
                            (flags & XTS_CF_EXACT_RADIUS) ? 1 : 0, (size_t)pool_cap);
    // pass 1: one tile per block by default (many short blocks keep the
    // last round of the grid short); a run of tiles per block on request
    const int ntiles = ca.tiles_w * ca.tiles_h;
    const int64_t want_blocks = (int64_t)ctx->cu_count * 4;
    const int64_t tpb = ctx->tune_tiles_per_block > 0 ? ctx->tune_tiles_per_block : 1;
    ca.tiles_per_block = (int)std::min<int64_t>(std::max<int64_t>(tpb, 1), ntiles);
    const int N4 = (int)((N + 3) / 4 * 4);
    const size_t tsz = f64 ? 8 : 4;
    const int64_t near_gx = std::min<int64_t>(std::max<int64_t>(1, 2 * want_blocks / S),
                                              div_up(plane, kNearPix));
    if (phase & kPhaseRows) {
        const size_t sm = (size_t)kNDPad * 8 +
                          (f64 ? (size_t)kNDPad * tsz + (size_t)N4 * 4 * tsz : 0);
        ca.rw_persist = ctx->tune_rw_blocks > 0 ? 1 : 0;
        const dim3 grid = ca.rw_persist
                              ? dim3((unsigned)std::min<int64_t>(ctx->tune_rw_blocks,
                                                                 (int64_t)S * ntiles))
                              : dim3((unsigned)div_up(ntiles, ca.tiles_per_block), (unsigned)S);
        TimedRegion tr(ctx, "cubature");
        if (!f64 && !ca.rw_persist && ctx->tune_tiles_per_block == 0 && !ctx->tune_rw_legacy) {
            // the tile-frame pass; the per-source tile blocks below on request
            // (tuning "rowweight_legacy", or their launch-shape keys)
            XTS_TRY(launch_rw_frames(ctx, ca, rw_frames_per_block(ctx, S, ntiles),
                                     ctx->tune_rw_frames_cu, ctx->stream));
        } else if (f64) {
            if (ca.rw_persist)
                rowweight_kernel<double, false, true><<<grid, kThreads, sm, ctx->stream>>>(ca);
            else
                rowweight_kernel<double><<<grid, kThreads, sm, ctx->stream>>>(ca);
        } else if (sur) {
            if (ca.rw_persist)
                rowweight_kernel<float, true, true><<<grid, kThreads, sm, ctx->stream>>>(ca);
            else
                rowweight_kernel<float, true><<<grid, kThreads, sm, ctx->stream>>>(ca);
        } else if (ca.rw_persist) {
            rowweight_kernel<float, false, true><<<grid, kThreads, sm, ctx->stream>>>(ca);
        } else {
            rowweight_kernel<float><<<grid, kThreads, sm, ctx->stream>>>(ca);
        }
        XTS_HIP(hipGetLastError());
    }
    // pass 1b: the surrogate's pixels, blocks stride over each source's list
    if (sur && (phase & kPhaseRows)) {
        TimedRegion tr(ctx, "surrogate");
        XTS_TRY(launch_sur(ctx, ca, want_blocks, 0, ctx->stream));
    }
    // pass 2a: blocks stride over each source's list
    if (phase & kPhaseRows) {
        const size_t sm = (size_t)kNDPad * tsz + (size_t)N4 * 4 * tsz + (size_t)N4 * 16 +
                          (size_t)kNearRows * kNearThreads * tsz;
        const dim3 grid((unsigned)near_gx, (unsigned)S);
        TimedRegion tr(ctx, "near");
        if (f64) {
            XTS_TRY(allow_lds((const void*)near_rows_kernel<double>, sm));
            near_rows_kernel<double><<<grid, kNearThreads, sm, ctx->stream>>>(ca);
        } else {
            XTS_TRY(allow_lds((const void*)near_rows_kernel<float>, sm));
            near_rows_kernel<float><<<grid, kNearThreads, sm, ctx->stream>>>(ca);
        }
        XTS_HIP(hipGetLastError());
    }
    if (full && (phase & kPhaseTable)) {
        TimedRegion tr(ctx, "factor");
        hist_integral_kernel<<<(unsigned)S, kHistThreads, 0, ctx->stream>>>(
            (int)E, ca.hist, upsf, ca.mapint, ig.geo, parint, ct->sep, fac,
            (flags & XTS_CF_EXACT_RADIUS) ? 1 : 0);
        XTS_HIP(hipGetLastError());
    }
    if (phase & kPhaseSum) {
        // the batch's maps x flux added to the summed model out [E][H W]
        ca.flux = flux;
        ca.sum = (float*)out;
        void* sf;
        XTS_TRY(slot_reserve(ctx, kSlotSumFac, (size_t)S * Ep * sizeof(float), &sf));
        ca.sum_fac = (float*)sf;
        TimedRegion tr(ctx, "sum");
        sum_factor_kernel<<<(unsigned)std::min<int64_t>(div_up((int64_t)S * Ep, 256), 1024), 256, 0,
                            ctx->stream>>>(ca);
        const unsigned nt = (unsigned)(ca.tiles_w * ca.tiles_h);
        const dim3 g48(nt, (unsigned)div_up(E, 48)), g112(nt, (unsigned)div_up(E, 112));
        const bool g4 = ctx->tune_sum_group != 8;
        if (E <= 48) {
            if (g4) sum_far_kernel<3, 4><<<g48, kThreads, 0, ctx->stream>>>(ca);
            else sum_far_kernel<3, 8><<<g48, kThreads, 0, ctx->stream>>>(ca);
        } else {
            if (g4) sum_far_kernel<7, 4><<<g112, kThreads, 0, ctx->stream>>>(ca);
            else sum_far_kernel<7, 8><<<g112, kThreads, 0, ctx->stream>>>(ca);
        }
        XTS_HIP(hipGetLastError());
        const size_t smw = (size_t)(kEmitThreads / 64) * kNDPad * sizeof(float);
        XTS_TRY(allow_lds((const void*)wide_sum_kernel<40>, smw));
        wide_sum_kernel<40><<<dim3(kWideBlocks, (unsigned)S, (unsigned)div_up(E, 40)), kEmitThreads,
                              smw, ctx->stream>>>(ca);
        XTS_HIP(hipGetLastError());
        const dim3 gn((unsigned)near_gx, (unsigned)S, (unsigned)div_up(E, 40));
        XTS_TRY((launch_near_emit<float, 40, float, true, true>(ctx->stream, ca, gn)));
        return XTS_OK;
    }
    if (!(phase & kPhaseEmit)) return XTS_OK;
    // pass 3: blocks of kEmitBlockPix pixels of one source, then kWideBlocks
    // blocks per source for the wide pixels (they depend only on the factors
    // and the near lists and write values no other pass writes, so they run
    // beside the far pixels' blocks)
    {
        ca.emit_gx = (int)div_up(plane, kEmitBlockPix);
        const dim3 grid32((unsigned)ca.emit_gx, (unsigned)S,
                          (unsigned)div_up(E, kEmitEC32));
        const dim3 grid64((unsigned)ca.emit_gx, (unsigned)S,
                          (unsigned)div_up(E, kEmitEC64));
        // the wide pixels by extra blocks of the same launch
        TimedRegion tr(ctx, "emit");
        if (f64) XTS_TRY((emit<double, kEmitEC64>(ctx->stream, ca, grid64, full)));
        else XTS_TRY((emit<float, kEmitEC32>(ctx->stream, ca, grid32, full)));
    }
    // pass 2b: listed pixels, scaled, written once
    {
        TimedRegion tr(ctx, "near_emit");
        if (f64) XTS_TRY((near_emit<double, 16>(ctx->stream, ca, near_gx, full)));
        else XTS_TRY((near_emit<float, 40>(ctx->stream, ca, near_gx, full)));
    }
    return XTS_OK;
}

// ------------------------------------------------------- no-allocation path
// The device memory of xts_point_src_model_map_wcs_dev_noalloc: every buffer
// a model map uses besides its inputs and output (pixel geometry, radii,
// constant tables, rule, PSF table copy, accumulators, row weights, near
// lists and pool) is carved from this array.  Each device holds its own
// copy from the library's load until the process ends.  Calls that overlap
// on one device share it and race.
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];
static int g_scratch_rc[kMaxDevices];
static bool g_scratch_done[kMaxDevices];

// Once per device, with that device current: the array's device address (the
// name is a host-side stand-in and must never be passed to a kernel).
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return XTS_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return g_scratch_rc[device];
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_rc[device] = e == hipSuccess && p ? XTS_OK : XTS_ERR_DEVICE;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_done[device] = true;
    if (e != hipSuccess) (void)hipGetLastError();   // leave no sticky error behind
    return g_scratch_rc[device];
}

int noalloc_scratch(int device, void** ptr, size_t* bytes) {
    XTS_CHECK(device >= 0 && device < kMaxDevices, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    XTS_CHECK(g_scratch_done[device], "no-allocation scratch not set up on device %d", device);
    if (g_scratch_rc[device] != XTS_OK)
        return fail(XTS_ERR_DEVICE, "hipGetSymbolAddress of the no-allocation scratch failed "
                    "on device %d", device);
    *ptr = g_scratch_ptr[device];
    *bytes = kNoAllocScratchBytes;
    return XTS_OK;
}

// Arena bytes model_map_device takes for one full model map: the requests it
// makes, in order, each rounded as slot_reserve rounds them.
static size_t model_map_arena_bytes(int64_t S, const xts_wcs* wcs, int64_t E, int64_t N,
                                    int32_t flags) {
    const bool f64 = (flags & XTS_COMPUTE_F64) != 0;
    const size_t H = (size_t)wcs->naxis1, W = (size_t)wcs->naxis2, plane = H * W;
    const size_t ncorn = (H + 1) * (W + 1), tsz = f64 ? 8 : 4;
    const size_t Ep = (size_t)((E + 3) / 4 * 4);
    const int64_t N4 = (N + 3) / 4 * 4;
    const size_t rec = ((size_t)N * sizeof(RuleUVW) + 255) & ~size_t(255);
    const int64_t pool_cap = std::min<int64_t>(4096 * kNearRows, ((int64_t)1 << 30) / S);
    size_t b = 0;
    b += arena_bytes(ncorn * 3 * 8 * 2 + plane * 8);                          // geometry
    b += arena_bytes((size_t)std::max<int64_t>(S, 1) * sizeof(SrcGeom));      // radii
    b += arena_bytes(const_tables_bytes());                                   // tables
    b += arena_bytes(rec + (size_t)kLeanArrays * N4 * sizeof(float));        // rule
    b += arena_bytes((size_t)S * kND * Ep * tsz);                             // PSF table
    b += arena_bytes((model_map_counter_doubles(S) + (size_t)S * kND + 2 * (size_t)S * E) *
                     8);                                                      // accumulators
    b += arena_bytes((size_t)S * plane * (f64 ? 32 : 16) + (size_t)S * plane * 20 + 512);
    b += arena_bytes((size_t)(pool_cap * S) * tsz);                          // near pool
    return b;
}

// The one-call device model map for the batched engine (xts_batch.hip).
int model_map_dev_call(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph, int64_t S,
                       const double* expo, const double* upsf, const double* parint, int64_t E,
                       const double* rule, int64_t N, int32_t flags, void* out, int phase) {
    XTS_CHECK(phase >= 1 && phase <= kPhaseAll, "bad phase %d", phase);
    return model_map_device(ctx, wcs, src_sph, S, expo, upsf, parint, E, rule, N, flags, out,
                            true, phase, nullptr, false);
}

static size_t out_bytes(int64_t S, const xts_wcs* wcs, int64_t E, int32_t flags) {
    return (size_t)S * wcs->naxis1 * wcs->naxis2 * E * ((flags & XTS_OUT_F32) ? 4 : 8);
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_point_src_model_map_wcs_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                    int64_t S, const double* expo, const double* upsf,
                                    const double* parint, int64_t E, const double* rule,
                                    int64_t N, int32_t flags, void* out) {
    XTS_TRY(activate(ctx));
    return model_map_device(ctx, wcs, src_sph, S, expo, upsf, parint, E, rule, N, flags, out, true);
}

int xts_point_src_summed_model_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                   int64_t S, const double* flux, const double* expo,
                                   const double* upsf, const double* parint, int64_t E,
                                   const double* rule, int64_t N, int32_t flags, float* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(S > 0 && S <= (1 << 24) && E > 0 && E <= 4096, "bad S=%lld or E=%lld",
              (long long)S, (long long)E);
    XTS_CHECK(src_sph && flux && expo && upsf && parint && out, "null device pointer");
    XTS_CHECK((flags & ~(XTS_SURROGATE | XTS_SURROGATE_NO_NET | XTS_SURROGATE_FP8 |
                         XTS_CF_EXACT_RADIUS)) == 0,
              "xts_point_src_summed_model_dev: flags 0x%x (only XTS_SURROGATE, "
              "XTS_SURROGATE_NO_NET, XTS_SURROGATE_FP8 and XTS_CF_EXACT_RADIUS apply)", flags);
    // the checks of the per-batch model maps, before anything is launched
    XTS_CHECK(!(flags & (XTS_SURROGATE_NO_NET | XTS_SURROGATE_FP8)) || (flags & XTS_SURROGATE),
              "XTS_SURROGATE_NO_NET and XTS_SURROGATE_FP8 need XTS_SURROGATE");
    XTS_CHECK(!(flags & XTS_SURROGATE_FP8) || !(flags & XTS_SURROGATE_NO_NET),
              "XTS_SURROGATE_FP8 needs the network");
    XTS_CHECK(!(flags & XTS_SURROGATE) || (flags & XTS_SURROGATE_NO_NET) || ctx->surrogate_ready,
              "XTS_SURROGATE: no network set (xts_ctx_set_surrogate)");
    const size_t plane = (size_t)wcs->naxis1 * wcs->naxis2;
    XTS_CHECK(plane <= (size_t)1 << 30, "image too large");
    const bool sur = (flags & XTS_SURROGATE) != 0;
    // sources per batch: the per-source working buffers (row weights, row
    // indices, near lists: 36 B per pixel; the surrogate's tile lists 32 B
    // more) within about 12 GB, unless tuned
    const double per_src = (double)plane * (36.0 + (sur ? 33.0 : 0.0)) + 401.0 * 4 * (E + 3);
    int64_t B = ctx->tune_sum_batch > 0 ? ctx->tune_sum_batch
                                        : (int64_t)std::max(1.0, std::floor(12e9 / per_src));
    B = std::max<int64_t>(1, std::min(B, S));
    XTS_CHECK(B * E < (1LL << 31), "batch of %lld sources x %lld energies too large",
              (long long)B, (long long)E);
    zero_kernel<<<(unsigned)std::min<size_t>(div_up(E * plane, 256), 65536), 256, 0, ctx->stream>>>(
        out, (size_t)E * plane);
    XTS_HIP(hipGetLastError());
    for (int64_t b0 = 0; b0 < S; b0 += B) {
        const int64_t n = std::min(B, S - b0);
        XTS_TRY(model_map_device(ctx, wcs, src_sph + 2 * b0, n, expo + b0 * E,
                                 upsf + (size_t)b0 * kND * E, parint + (size_t)b0 * kND * E, E,
                                 rule, N, flags | XTS_LAYOUT_SEHW | XTS_OUT_F32, out, true,
                                 kPhaseRows | kPhaseTable | kPhaseSum, flux + b0));
    }
    return XTS_OK;
}

int xts_point_src_model_map_wcs_dev_noalloc(xts_ctx* ctx, const xts_wcs* wcs,
                                            const double* src_sph, int64_t S,
                                            const double* expo, const double* upsf,
                                            const double* parint, int64_t E, const double* rule,
                                            int64_t N, int32_t flags, void* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(S > 0 && S <= (1 << 24) && E > 0 && E <= 4096 && S * E < (1LL << 31),
              "bad S=%lld or E=%lld", (long long)S, (long long)E);
    XTS_CHECK(rule == nullptr || (N > 0 && N <= 1024), "rule size %lld out of range (1..1024)",
              (long long)N);
    XTS_CHECK((int64_t)wcs->naxis1 * wcs->naxis2 <= (int64_t)1 << 30, "image too large");
    XTS_CHECK(!(flags & (XTS_SURROGATE | XTS_SURROGATE_NO_NET | XTS_SURROGATE_FP8)),
              "xts_point_src_model_map_wcs_dev_noalloc: the surrogate is not offered here");
    void* base;
    size_t cap;
    XTS_TRY(noalloc_scratch(ctx->device, &base, &cap));
    const size_t need = model_map_arena_bytes(S, wcs, E, rule ? N : 12, flags);
    XTS_CHECK(need <= cap,
              "xts_point_src_model_map_wcs_dev_noalloc: %lld sources x %lld x %lld pixels x "
              "%lld energies need %zu MiB of scratch, more than its %zu MiB",
              (long long)S, (long long)wcs->naxis1, (long long)wcs->naxis2, (long long)E,
              (need + (1 << 20) - 1) >> 20, cap >> 20);
    struct ArenaScope {
        xts_ctx* c;
        ~ArenaScope() { c->arena.base = nullptr; c->arena.cap = c->arena.used = 0; }
    } scope{ctx};
    ctx->arena.base = (char*)base;
    ctx->arena.cap = cap;
    ctx->arena.used = 0;
    return model_map_device(ctx, wcs, src_sph, S, expo, upsf, parint, E, rule, N, flags, out, true);
}

int xts_point_src_model_map_rows_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                     int64_t S, int64_t E, const double* rule, int64_t N,
                                     int32_t flags) {
    XTS_TRY(activate(ctx));
    ctx->pending_rows.valid = false;
    XTS_TRY(model_map_device(ctx, wcs, src_sph, S, nullptr, nullptr, nullptr, E, rule, N, flags,
                             nullptr, true, kPhaseRows));
    auto& p = ctx->pending_rows;
    p.wcs = *wcs;
    p.src = src_sph;
    p.rule = rule;
    p.S = S; p.E = E; p.N = N;
    p.flags = flags;
    p.valid = true;
    return XTS_OK;
}

int xts_point_src_model_map_finish_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                       int64_t S, const double* expo, const double* upsf,
                                       const double* parint, int64_t E, const double* rule,
                                       int64_t N, int32_t flags, void* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(wcs != nullptr, "null wcs");
    auto& p = ctx->pending_rows;
    // the rule by content: the row-weight half uploaded it, and a wrapper may
    // pass an equal copy
    const bool same_rule =
        rule ? (p.rule != nullptr && ctx->rule.host.size() == (size_t)(3 * N) &&
                std::equal(rule, rule + 3 * N, ctx->rule.host.begin()))
             : p.rule == nullptr;
    XTS_CHECK(p.valid && std::memcmp(&p.wcs, wcs, sizeof(xts_wcs)) == 0 && p.src == src_sph &&
                  same_rule && p.S == S && p.E == E && p.N == N && p.flags == flags,
              "xts_point_src_model_map_finish_dev: no row-weight half with these arguments "
              "(xts_point_src_model_map_rows_dev)");
    p.valid = false;
    return model_map_device(ctx, wcs, src_sph, S, expo, upsf, parint, E, rule, N, flags, out,
                            true, kPhaseFinish);
}

int xts_point_src_model_map_wcs(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                int64_t S, const double* expo, const double* upsf,
                                const double* parint, int64_t E, const double* rule, int64_t N,
                                int32_t flags, void* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(src_sph && expo && upsf && parint && out, "null host pointer");
    XTS_CHECK(S > 0 && E > 0, "bad sizes");
    void *dsrc, *dexp, *dupsf, *dpar, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, src_sph, S * 2 * 8, &dsrc));
    XTS_TRY(to_device(ctx, kSlotIn1, expo, S * E * 8, &dexp));
    XTS_TRY(to_device(ctx, kSlotIn2, upsf, (size_t)S * kND * E * 8, &dupsf));
    XTS_TRY(to_device(ctx, kSlotIn3, parint, (size_t)S * kND * E * 8, &dpar));
    const size_t ob = out_bytes(S, wcs, E, flags);
    XTS_TRY(slot_reserve(ctx, kSlotOut0, ob, &dout));
    XTS_TRY(model_map_device(ctx, wcs, (double*)dsrc, S, (double*)dexp, (double*)dupsf,
                             (double*)dpar, E, rule, N, flags, dout, true));
    XTS_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_cubature_convolve_map_psf(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                  int64_t S, const double* upsf, int64_t E, const double* rule,
                                  int64_t N, int32_t flags, void* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(src_sph && upsf && out && S > 0 && E > 0, "bad arguments");
    void *dsrc, *dupsf, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, src_sph, S * 2 * 8, &dsrc));
    XTS_TRY(to_device(ctx, kSlotIn2, upsf, (size_t)S * kND * E * 8, &dupsf));
    const size_t ob = out_bytes(S, wcs, E, flags);
    XTS_TRY(slot_reserve(ctx, kSlotOut0, ob, &dout));
    XTS_TRY(model_map_device(ctx, wcs, (double*)dsrc, S, nullptr, (double*)dupsf, nullptr, E,
                             rule, N, flags, dout, false));
    XTS_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_solid_angle(xts_ctx* ctx, const xts_wcs* wcs, double* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    ctx->pending_rows.valid = false;   // rebuilds the geometry slots
    XTS_CHECK(out != nullptr, "null out");
    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, nullptr, 0, &ig));
    XTS_HIP(hipMemcpyAsync(out, ig.omega, (size_t)wcs->naxis1 * wcs->naxis2 * 8,
                           hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_psf_boundary_radius(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph, int64_t S,
                            double* radius, uint8_t* in_fov) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(src_sph && radius && in_fov && S > 0, "bad arguments");
    ctx->pending_rows.valid = false;   // rebuilds the geometry slots
    void* dsrc;
    XTS_TRY(to_device(ctx, kSlotIn0, src_sph, S * 2 * 8, &dsrc));
    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, (double*)dsrc, (int)S, &ig));
    std::vector<SrcGeom> g((size_t)S);
    XTS_HIP(hipMemcpyAsync(g.data(), ig.geo, S * sizeof(SrcGeom), hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < S; ++s) {
        radius[s] = g[s].radius_deg;
        in_fov[s] = (uint8_t)g[s].in_fov;
    }
    return XTS_OK;
}

int xts_map_integral(xts_ctx* ctx, const xts_wcs* wcs, const double* map, const double* src_sph,
                     int64_t S, int64_t E, const double* radius, const uint8_t* in_fov,
                     double* out) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(map && src_sph && radius && in_fov && out && S > 0 && E > 0, "bad arguments");
    ctx->pending_rows.valid = false;   // rebuilds the geometry slots
    const int H = (int)wcs->naxis1, W = (int)wcs->naxis2;
    void *dsrc, *dmap, *dg;
    XTS_TRY(to_device(ctx, kSlotIn0, src_sph, S * 2 * 8, &dsrc));
    XTS_TRY(to_device(ctx, kSlotIn1, map, (size_t)S * H * W * E * 8, &dmap));
    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, nullptr, 0, &ig));
    std::vector<SrcGeom> g((size_t)S);
    for (int64_t s = 0; s < S; ++s) {
        sph2dir(src_sph[2 * s], src_sph[2 * s + 1], g[s].dir);
        g[s].radius_deg = radius[s];
        g[s].in_fov = in_fov[s];
    }
    XTS_TRY(to_device(ctx, kSlotIn2, g.data(), S * sizeof(SrcGeom), &dg));
    void* dout;
    XTS_TRY(slot_reserve(ctx, kSlotOut0, S * E * 8, &dout));
    map_integral_kernel<<<(unsigned)S, 64, 0, ctx->stream>>>(H, W, (int)E, (double*)dmap, ig.centres,
                                                             (SrcGeom*)dg, (double*)dout);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, S * E * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_map_correction_factor(xts_ctx* ctx, const double* invmi, const double* radius,
                              const uint8_t* in_fov, const double* upsf, const double* parint,
                              int64_t S, int64_t E, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(invmi && radius && in_fov && upsf && parint && out && S > 0 && E > 0,
              "bad arguments");
    std::vector<SrcGeom> g((size_t)S);
    for (int64_t s = 0; s < S; ++s) {
        g[s].radius_deg = radius[s];
        g[s].in_fov = in_fov[s];
    }
    void *dg, *dup, *dpar, *dinv, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, g.data(), S * sizeof(SrcGeom), &dg));
    XTS_TRY(to_device(ctx, kSlotIn1, upsf, (size_t)S * kND * E * 8, &dup));
    XTS_TRY(to_device(ctx, kSlotIn2, parint, (size_t)S * kND * E * 8, &dpar));
    XTS_TRY(to_device(ctx, kSlotIn3, invmi, S * E * 8, &dinv));
    const xts_const_tables* ct;
    XTS_TRY(const_tables(ctx, &ct));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, S * E * 8, &dout));
    factor_kernel<<<div_up(S * E, 256), 256, 0, ctx->stream>>>(
        (int)S, (int)E, (SrcGeom*)dg, nullptr, (double*)dinv, (double*)dup, (double*)dpar, ct->sep,
        nullptr, (double*)dout, nullptr);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, S * E * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_scale_map(xts_ctx* ctx, double* map, int64_t S, int64_t H, int64_t W, int64_t E,
                  const double* factor, int32_t kind) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(map && factor && S > 0 && H > 0 && W > 0 && E > 0, "bad arguments");
    XTS_CHECK(kind == XTS_SCALE_PIXEL || kind == XTS_SCALE_SRC_ENERGY, "bad kind %d", kind);
    const size_t n = (size_t)S * H * W * E;
    void *dmap, *df;
    XTS_TRY(to_device(ctx, kSlotIn0, map, n * 8, &dmap));
    const size_t fn = kind == XTS_SCALE_PIXEL ? (size_t)H * W : (size_t)S * E;
    XTS_TRY(to_device(ctx, kSlotIn1, factor, fn * 8, &df));
    const unsigned blocks = std::min<unsigned>(div_up((int64_t)n, 256), 65536);
    if (kind == XTS_SCALE_PIXEL)
        scale_pixel_kernel<<<blocks, 256, 0, ctx->stream>>>((double*)dmap, (size_t)S * H * W,
                                                            (int)(H * W), (int)E, (double*)df);
    else
        scale_se_kernel<<<blocks, 256, 0, ctx->stream>>>((double*)dmap, (size_t)H * W, (int)S,
                                                         (int)E, (double*)df);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(map, dmap, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_pix2sph(xts_ctx* ctx, const xts_wcs* wcs, const double* pix, int64_t n, double* sph) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(pix && sph && n > 0, "bad arguments");
    void *din, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, pix, n * 16, &din));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 16, &dout));
    pix2sph_kernel<<<std::min<unsigned>(div_up(n, 256), 4096), 256, 0, ctx->stream>>>(
        wcs_setup(*wcs), n, (double*)din, (double*)dout, 0);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(sph, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_sph2pix(xts_ctx* ctx, const xts_wcs* wcs, const double* sph, int64_t n, double* pix) {
    XTS_TRY(activate(ctx));
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(pix && sph && n > 0, "bad arguments");
    void *din, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, sph, n * 16, &din));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, n * 16, &dout));
    pix2sph_kernel<<<std::min<unsigned>(div_up(n, 256), 4096), 256, 0, ctx->stream>>>(
        wcs_setup(*wcs), n, (double*)din, (double*)dout, 1);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(pix, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

}  // extern "C"
