// xts_modelmap.hip -- (c) point-source model map on gfx950.
//
// Pipeline for one image (model_map.point_src_model_map_wcs.cxx:6-45):
//   1. geometry_kernel   FP64 pixel-corner and pixel-centre unit vectors of the
//                        WCS grid and the per-pixel solid angle (once per image)
//   2. radius_kernel     in-FOV flag and boundary radius per source
//   3. cubature_kernel   per (source, pixel tile, energy chunk): PSF table rows
//                        staged in LDS, cubature of the table over each pixel,
//                        x solid angle, stores, and the in-radius partial sums
//                        of the map integral
//   4. factor_kernel     correction factor x exposure per (source, energy)
//   5. scale_kernel      in-place multiply of every map value by its factor
#include <algorithm>
#include <cstring>
#include <vector>

#include "xts_internal.hpp"

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
    double chord_lo, chord_hi;  // bracket of 2 sin(radius / 2)
    int in_fov;
    int pad;
};

__global__ void radius_kernel(Wcs w, int H, int W, int S, const double* __restrict__ src_sph,
                              const double* __restrict__ corners, SrcGeom* __restrict__ geo) {
    const int s = blockIdx.x;
    if (s >= S) return;
    __shared__ double red[256];
    double src[3];
    sph2dir(src_sph[2 * s], src_sph[2 * s + 1], src);
    const int nb = 2 * (H + 1) + 2 * (W + 1);
    double mn = 360.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        int ci, cj;
        if (k < H + 1) { ci = k; cj = 0; }
        else if (k < 2 * (H + 1)) { ci = k - (H + 1); cj = W; }
        else if (k < 2 * (H + 1) + W + 1) { ci = 0; cj = k - 2 * (H + 1); }
        else { ci = H; cj = k - 2 * (H + 1) - (W + 1); }
        const double d = kR2D * dir_diff(src, corners + 3 * (ci * (W + 1) + cj));
        mn = d < mn ? d : mn;
    }
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double p1, p2;
        wcs_sph2pix(w, src_sph[2 * s], src_sph[2 * s + 1], &p1, &p2);
        const double buf = 3.5;
        SrcGeom g;
        g.dir[0] = src[0]; g.dir[1] = src[1]; g.dir[2] = src[2];
        g.radius_deg = red[0];
        g.in_fov = p1 > buf && p1 < (H - buf) && p2 > buf && p2 < (W - buf);
        const double c = 2.0 * sin(0.5 * kD2R * red[0]);
        g.chord_lo = c * (1.0 - 1e-12);
        g.chord_hi = c * (1.0 + 1e-12);
        g.pad = 0;
        geo[s] = g;
    }
}

// ------------------------------------------------------------------ cubature
// Rule point constants: bilinear corner weights b0..b3 of
// points_weights.hpp:37-51 (corner order LL, LR, UR, UL) and the weight w/4.
struct RulePoint {
    double b[4];
    double wq;
};
struct RulePoint32 {
    float b[4];
    float wq;
    float pad[3];
};

constexpr int kTileW = 16, kTileH = 16;   // pixels per block pass (256 threads)
constexpr int kThreads = kTileW * kTileH;
constexpr int kWin = 8;                   // node-weight window per pixel
constexpr int kWinStride = kWin + 1;      // odd stride: conflict-free lanes

struct CubArgs {
    const double* corners;
    const double* centres;
    const double* omega;       // may be null (raw cubature)
    const SrcGeom* geo;
    const void* lut;           // [S][401][Ep] of T
    const RulePoint* rule;
    const RulePoint32* rule32;
    const float* inv_sep;      // [401] 1 / sep[i] (deg), FP32 table
    const double* inv_sep64;   // [401] FP64 table
    double* mapint;            // [S][E] or null
    void* out;
    int S, H, W, E, Ep, N;
    int tiles_w, tiles_h, tiles_per_block;
    int layout_sehw, out_f32;
    double rho_scale;          // max over points of sum_k |b_k|
};

// upper node index of the reference's logsep lookup
// (convolve_pixel_psf_logsep.hpp:46-53): 2 below 1e-4 deg, clamped to 400.
template <typename T>
__device__ inline int sep_index(T s) {
    if (s < T(1e-4)) return 2;
    const T j = T(1) + T(kRecipStep) * (T(kLn1e4) + log(s));
    int i = 1 + (int)j;
    return i > 400 ? 400 : i;
}

__device__ inline float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994530942f; }
__device__ inline double fast_log(double x) { return log(x); }
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fast_sqrt(double x) { return sqrt(x); }

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

// T: cubature arithmetic type; EC: energies per block (multiple of the
// vector width).  Dynamic LDS: LUT rows [401][EC] of T, the per-thread
// node-weight windows [kThreads][kWinStride] (reused for the block
// reduction), 1/sep [kNDPad], then the rule (b0..b3 vectors, weights).
//
// Per pixel the separations of all rule points lie in [c0 - rho, c0 + rho]
// (c0: chord to the pixel's bilinear centre, rho: largest corner offset
// from it), so the LUT rows they touch are known before the point loop.
// Pixels whose points touch at most four consecutive rows accumulate the
// per-row interpolation weights in four registers (hat functions of the
// fractional node position); the energy loop then reads four LUT rows.
// Other pixels (near the source) add each point's two rows directly.
template <typename T, int EC>
__global__ void __launch_bounds__(kThreads)
cubature_kernel(CubArgs a) {
    using V = typename Vec<T>::type;
    constexpr int VN = Vec<T>::n;
    constexpr int NV = EC / VN;
    using V4 = typename Vec4<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lut_s = reinterpret_cast<T*>(smem_raw);
    T* win_s = lut_s + kND * EC;
    T* isep_s = win_s + kThreads * kWinStride;
    V4* rb_s = reinterpret_cast<V4*>(isep_s + kNDPad);
    T* rw_s = reinterpret_cast<T*>(rb_s + a.N);

    const int s = blockIdx.y;
    const int e0 = blockIdx.z * EC;
    const int ne = min(EC, a.E - e0);
    const int tid = threadIdx.x;
    const int ntiles = a.tiles_w * a.tiles_h;
    const int t_begin = blockIdx.x * a.tiles_per_block;
    const int t_end = min(ntiles, t_begin + a.tiles_per_block);
    if (t_begin >= ntiles) return;

    // ---- stage this source's PSF rows (energies e0 .. e0+EC), 1/sep and the rule
    const T* lut_g = reinterpret_cast<const T*>(a.lut) + (size_t)s * kND * a.Ep + e0;
    for (int k = tid; k < kND * NV; k += kThreads) {
        const int d = k / NV, v = k % NV;
        V val;
        const T* src = lut_g + (size_t)d * a.Ep + v * VN;
        if (e0 + v * VN + VN <= a.Ep) val = *reinterpret_cast<const V*>(src);
        else { T tmp[VN] = {}; val = *reinterpret_cast<V*>(tmp); }
        reinterpret_cast<V*>(lut_s)[d * NV + v] = val;
    }
    for (int k = tid; k < kND; k += kThreads)
        isep_s[k] = sizeof(T) == 4 ? (T)a.inv_sep[k] : (T)a.inv_sep64[k];
    for (int p = tid; p < a.N; p += kThreads) {
        V4 b;
        if (sizeof(T) == 4) {
            const RulePoint32 rp = a.rule32[p];
            b = V4{(T)rp.b[0], (T)rp.b[1], (T)rp.b[2], (T)rp.b[3]};
            rw_s[p] = (T)rp.wq;
        } else {
            const RulePoint rp = a.rule[p];
            b = V4{(T)rp.b[0], (T)rp.b[1], (T)rp.b[2], (T)rp.b[3]};
            rw_s[p] = (T)rp.wq;
        }
        rb_s[p] = b;
    }
    __syncthreads();

    const SrcGeom g = a.geo[s];
    const double sx = g.dir[0], sy = g.dir[1], sz = g.dir[2];
    const int lx = tid % kTileW, ly = tid / kTileW;
    const int Wp1 = a.W + 1;

    // per-thread running sums of the map integral (pixels within the radius)
    T isum[EC];
#pragma unroll
    for (int e = 0; e < EC; ++e) isum[e] = T(0);

    for (int t = t_begin; t < t_end; ++t) {
        const int th = t / a.tiles_w, tw = t % a.tiles_w;
        const int h = th * kTileH + ly, w = tw * kTileW + lx;
        const bool valid = h < a.H && w < a.W;
        // out-of-image lanes compute an in-image pixel and store nothing, so
        // the point loop runs in uniform control flow
        const int hc = min(h, a.H - 1), wc = min(w, a.W - 1);

        // corner deltas corner_k - source in FP64, then T (LL LR UR UL)
        const int c00 = hc * Wp1 + wc;
        const int cidx[4] = {c00, c00 + 1, c00 + Wp1 + 1, c00 + Wp1};
        T D[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double* c = a.corners + 3 * cidx[k];
            D[k][0] = (T)(c[0] - sx);
            D[k][1] = (T)(c[1] - sy);
            D[k][2] = (T)(c[2] - sz);
        }
        const T cx = T(0.25) * (D[0][0] + D[1][0] + D[2][0] + D[3][0]);
        const T cy = T(0.25) * (D[0][1] + D[1][1] + D[2][1] + D[3][1]);
        const T cz = T(0.25) * (D[0][2] + D[1][2] + D[2][2] + D[3][2]);
        const T c0 = fast_sqrt(cx * cx + cy * cy + cz * cz);
        T rho = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const T ex = D[k][0] - cx, ey = D[k][1] - cy, ez = D[k][2] - cz;
            rho = fmax(rho, fast_sqrt(ex * ex + ey * ey + ez * ez));
        }
        rho = rho * (T)a.rho_scale * T(1.0001) + T(1e-6) * c0;
        const T slo = T(2.0 * kR2D) * asin_pos(T(0.5) * fmax(c0 - rho, T(0)));
        const T shi = T(2.0 * kR2D) * asin_pos(fmin(T(0.5) * (c0 + rho), T(1)));
        const int ilo = sep_index_fast(slo), ihi = sep_index_fast(shi);
        // rows touched: ilo-1 .. ihi; inside the table with t in [0, 1]
        const bool fast = (ihi - ilo) <= 2 && slo >= T(1e-4) && shi < T(69.9);
        const int r0 = min(max(ilo - 1, 0), kND - 4);

        V acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = V{};
        T W0 = T(0), W1 = T(0), W2 = T(0), W3 = T(0);
        const T fr0 = (T)(r0 + 1);

#pragma unroll 4
        for (int p = 0; p < a.N; ++p) {
            const V4 bv = rb_s[p];
            const T wq = rw_s[p];
            const T dx = bv.x * D[0][0] + bv.y * D[1][0] + bv.z * D[2][0] + bv.w * D[3][0];
            const T dy = bv.x * D[0][1] + bv.y * D[1][1] + bv.z * D[2][1] + bv.w * D[3][1];
            const T dz = bv.x * D[0][2] + bv.y * D[1][2] + bv.z * D[2][2] + bv.w * D[3][2];
            const T ch = fast_sqrt(dx * dx + dy * dy + dz * dz);
            const T sd = T(2.0 * kR2D) * asin_pos(fmin(T(0.5) * ch, T(1)));
            const int i = sep_index_fast(sd);
            // t = (s - sep[i-1]) / (sep[i] - sep[i-1]) = redm1 (s / sep[i-1] - 1)
            const T tt = T(kRedm1) * (sd * isep_s[i - 1] - T(1));
            if (fast) {
                // fractional node position relative to row r0
                const T xx = ((T)i - fr0) + tt;
                W0 += wq * fmax(T(0), T(1) - fabs(xx));
                W1 += wq * fmax(T(0), T(1) - fabs(xx - T(1)));
                W2 += wq * fmax(T(0), T(1) - fabs(xx - T(2)));
                W3 += wq * fmax(T(0), T(1) - fabs(xx - T(3)));
            } else {
                const T wa = wq * (T(1) - tt), wb = wq * tt;
                const V* L0 = reinterpret_cast<const V*>(lut_s) + (i - 1) * NV;
                const V* L1 = L0 + NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    vfma(acc[v], wa, L0[v]);
                    vfma(acc[v], wb, L1[v]);
                }
            }
        }
        if (fast) {
            const V* L = reinterpret_cast<const V*>(lut_s) + r0 * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                vfma(acc[v], W0, L[v]);
                vfma(acc[v], W1, L[NV + v]);
                vfma(acc[v], W2, L[2 * NV + v]);
                vfma(acc[v], W3, L[3 * NV + v]);
            }
        }
        T om = T(1);
        if (a.omega) om = (T)a.omega[hc * a.W + wc];
        bool inrad = false;
        if (a.mapint && g.in_fov && valid) {
            const double* c = a.centres + 3 * (hc * Wp1 + wc);
            const double ex = g.dir[0] - c[0], ey = g.dir[1] - c[1], ez = g.dir[2] - c[2];
            const double chd = sqrt(ex * ex + ey * ey + ez * ez);
            inrad = chd < g.chord_lo ? true
                  : (chd > g.chord_hi ? false
                                      : (2.0 * asin(0.5 * chd) <= kD2R * g.radius_deg));
        }
        // ---- epilogue: x solid angle, store, in-radius partial sums
        if (valid) {
            const size_t plane = (size_t)a.H * a.W;
            const size_t base = a.layout_sehw
                                    ? ((size_t)s * a.E + e0) * plane + (size_t)h * a.W + w
                                    : (((size_t)s * a.H + h) * a.W + w) * a.E + e0;
            const size_t step = a.layout_sehw ? plane : 1;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
#pragma unroll
                for (int j = 0; j < VN; ++j) {
                    const int e = v * VN + j;
                    if (e < ne) {
                        const T val = reinterpret_cast<const T*>(&acc[v])[j] * om;
                        if (a.out_f32) reinterpret_cast<float*>(a.out)[base + e * step] = (float)val;
                        else reinterpret_cast<double*>(a.out)[base + e * step] = (double)val;
                        if (inrad) isum[e] += val;
                    }
                }
            }
        }
    }

    // ---- map integral: reduce the per-thread sums over the block, then one
    // FP64 atomic per (source, energy) per block.
    if (a.mapint && g.in_fov) {
        __syncthreads();
        double* red = reinterpret_cast<double*>(win_s);  // reuse the window space
        const int wave = tid / 64, lane = tid % 64;
        const int nwaves = kThreads / 64;
#pragma unroll
        for (int e = 0; e < EC; ++e) {
            if (e < ne) {
                T v = isum[e];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) red[wave * EC + e] = (double)v;
            }
        }
        __syncthreads();
        if (tid < ne) {
            double tot = 0.0;
            for (int k = 0; k < nwaves; ++k) tot += red[k * EC + tid];
            atomicAdd(&a.mapint[(size_t)s * a.E + e0 + tid], tot);
        }
    }
}

// ------------------------------------------------------------------ factors
// model_map.map_integral.cxx:41-44 (reciprocal), map_correction_factor.cxx:
// 24-50 and scale_map_by_exposure: f[s, e] = exposure x correction.
__global__ void factor_kernel(int S, int E, const SrcGeom* __restrict__ geo,
                              const double* __restrict__ mapint,
                              const double* __restrict__ invmi_in, const double* __restrict__ upsf,
                              const double* __restrict__ parint, const double* __restrict__ sep,
                              const double* __restrict__ expo, double* __restrict__ cf_out,
                              double* __restrict__ invmi_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
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
    double cf = 1.0;
    if (g.in_fov) {
        const double r = g.radius_deg;
        int d = r < 1.0 ? 1 : (int)floor(1.0 + log(1e4 * r) * kRecipStep);
        if (d >= kND - 2) d = kND - 2;
        const double t1 = sep[d] * kD2R, t2 = sep[d + 1] * kD2R;
        const double y1 = kTwoPi * upsf[((size_t)s * kND + d) * E + e] * sin(t1);
        const double y2 = kTwoPi * upsf[((size_t)s * kND + d + 1) * E + e] * sin(t2);
        const double m = (y2 - y1) / (t2 - t1);
        const double b = y1 - t1 * m;
        const double v = 0.5 * m * (t2 * t2 - t1 * t1) + b * (t2 - t1);
        double integ = v + parint[((size_t)s * kND + d) * E + e];
        integ = integ > 1.0 ? 1.0 : integ;
        cf = integ * invmi;
    }
    if (cf_out) cf_out[k] = cf;
    if (invmi_out) invmi_out[k] = invmi;
}

// map[s, ...] *= expo[s, e] then *= cf[s, e] (two roundings, as the
// reference applies them in two passes).
template <typename OT>
__global__ void scale_kernel(OT* __restrict__ out, int S, int E, size_t plane, int layout_sehw,
                             const double* __restrict__ expo, const double* __restrict__ cf) {
    const size_t n = (size_t)S * E * plane;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        size_t se;
        if (layout_sehw) se = k / plane;
        else { const size_t s = k / (plane * E); se = s * E + (k % E); }
        if (sizeof(OT) == 8) {
            double v = (double)out[k];
            v *= expo[se];
            v *= cf[se];
            out[k] = (OT)v;
        } else {
            out[k] = (OT)((double)out[k] * (expo[se] * cf[se]));
        }
    }
}

// [S, E, H, W]: blockIdx.y is the (s, e) plane, factor read once per block
template <typename OT>
__global__ void scale_planes_kernel(OT* __restrict__ out, size_t plane,
                                    const double* __restrict__ expo, const double* __restrict__ cf) {
    const size_t se = blockIdx.y;
    OT* p = out + se * plane;
    if (sizeof(OT) == 8) {
        const double f0 = expo[se], f1 = cf[se];
        for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < plane;
             k += (size_t)gridDim.x * blockDim.x) {
            double v = (double)p[k];
            v *= f0;
            v *= f1;
            p[k] = (OT)v;
        }
    } else {
        const float f = (float)(expo[se] * cf[se]);
        const size_t n4 = plane / 4;
        float4* p4 = reinterpret_cast<float4*>(p);
        const bool aligned = (reinterpret_cast<uintptr_t>(p) & 15) == 0;
        if (aligned) {
            for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4;
                 k += (size_t)gridDim.x * blockDim.x) {
                float4 v = p4[k];
                v.x *= f; v.y *= f; v.z *= f; v.w *= f;
                p4[k] = v;
            }
            for (size_t k = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < plane;
                 k += (size_t)gridDim.x * blockDim.x)
                p[k] = (OT)((float)p[k] * f);
        } else {
            for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < plane;
                 k += (size_t)gridDim.x * blockDim.x)
                p[k] = (OT)((float)p[k] * f);
        }
    }
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

__global__ void zero_kernel(double* __restrict__ p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
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
                          ImageGeom* ig) {
    const int H = (int)wcs->naxis1, W = (int)wcs->naxis2;
    const size_t ncorn = (size_t)(H + 1) * (W + 1);
    const size_t bytes = ncorn * 3 * 8 * 2 + (size_t)H * W * 8 + (size_t)std::max(S, 1) * sizeof(SrcGeom) + 1024;
    void* base;
    XTS_TRY(slot_reserve(ctx, kSlotGeom, bytes, &base));
    char* p = (char*)base;
    ig->corners = (double*)p; p += ncorn * 3 * 8;
    ig->centres = (double*)p; p += ncorn * 3 * 8;
    ig->omega = (double*)p; p += (size_t)H * W * 8;
    p = (char*)(((uintptr_t)p + 255) & ~uintptr_t(255));
    ig->geo = (SrcGeom*)p;
    const Wcs w = wcs_setup(*wcs);
    geometry_kernel<<<std::min<unsigned>(div_up(ncorn, 256), 4096), 256, 0, ctx->stream>>>(
        w, H, W, ig->corners, ig->centres);
    solid_angle_kernel<<<std::min<unsigned>(div_up((int64_t)H * W, 256), 4096), 256, 0, ctx->stream>>>(
        H, W, ig->centres, ig->omega);
    if (S > 0 && src_sph_dev)
        radius_kernel<<<S, 256, 0, ctx->stream>>>(w, H, W, S, src_sph_dev, ig->corners, ig->geo);
    XTS_HIP(hipGetLastError());
    return XTS_OK;
}

// Separation tables (FP64 and FP32 inverse) in the kSlotWork3 slot.
struct SepTables {
    double* sep;
    double* inv64;
    float* inv32;
};

static int sep_tables(xts_ctx* ctx, SepTables* t) {
    static std::vector<char> host;
    if (host.empty()) {
        host.resize(kND * (8 + 8 + 4));
        double* sep = (double*)host.data();
        double* inv = sep + kND;
        float* inv32 = (float*)(inv + kND);
        const double xstep = std::log(70.0 / 1e-4) / 399.0;
        sep[0] = 0.0;
        for (int i = 0; i < 400; ++i) sep[i + 1] = 1e-4 * std::exp(i * xstep);
        for (int i = 0; i < kND; ++i) {
            inv[i] = sep[i] > 0 ? 1.0 / sep[i] : 0.0;
            inv32[i] = (float)inv[i];
        }
    }
    void* d;
    XTS_TRY(to_device(ctx, kSlotWork3, host.data(), host.size(), &d));
    t->sep = (double*)d;
    t->inv64 = t->sep + kND;
    t->inv32 = (float*)(t->inv64 + kND);
    return XTS_OK;
}

// FP32 copy of the PSF table, energy axis padded to a multiple of 4.
__global__ void lut_convert_kernel(const double* __restrict__ upsf, int S, int E, int Ep,
                                   float* __restrict__ out32, double* __restrict__ out64) {
    const size_t n = (size_t)S * kND * Ep;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int e = (int)(k % Ep);
        const size_t sd = k / Ep;
        const double v = e < E ? upsf[sd * E + e] : 0.0;
        if (out32) out32[k] = (float)v;
        if (out64) out64[k] = v;
    }
}

static int upload_rule(xts_ctx* ctx, const double* rule, int64_t N, RulePoint** dev,
                       RulePoint32** dev32, double* rho_scale) {
    const double* r = rule ? rule : kSquareDeg7;
    if (!rule) N = 12;
    XTS_CHECK(N > 0 && N <= 1024, "rule size %lld out of range (1..1024)", (long long)N);
    std::vector<RulePoint> rp((size_t)N);
    double rs = 1.0;
    for (int64_t p = 0; p < N; ++p) {
        const double u = r[3 * p], v = r[3 * p + 1];
        rp[p].b[0] = 0.25 * (1.0 - u) * (1.0 - v);
        rp[p].b[1] = 0.25 * (1.0 + u) * (1.0 - v);
        rp[p].b[2] = 0.25 * (1.0 + u) * (1.0 + v);
        rp[p].b[3] = 0.25 * (1.0 - u) * (1.0 + v);
        rp[p].wq = 0.25 * r[3 * p + 2];
        rs = std::max(rs, std::fabs(rp[p].b[0]) + std::fabs(rp[p].b[1]) +
                              std::fabs(rp[p].b[2]) + std::fabs(rp[p].b[3]));
    }
    std::vector<char> buf(rp.size() * (sizeof(RulePoint) + sizeof(RulePoint32)));
    memcpy(buf.data(), rp.data(), rp.size() * sizeof(RulePoint));
    RulePoint32* r32 = (RulePoint32*)(buf.data() + rp.size() * sizeof(RulePoint));
    for (size_t p = 0; p < rp.size(); ++p) {
        for (int k = 0; k < 4; ++k) r32[p].b[k] = (float)rp[p].b[k];
        r32[p].wq = (float)rp[p].wq;
        r32[p].pad[0] = r32[p].pad[1] = r32[p].pad[2] = 0.f;
    }
    void* d;
    XTS_TRY(to_device(ctx, kSlotWork2, buf.data(), buf.size(), &d));
    *dev = (RulePoint*)d;
    *dev32 = (RulePoint32*)((char*)d + rp.size() * sizeof(RulePoint));
    *rho_scale = rs;
    return XTS_OK;
}

// The device pipeline shared by the host and _dev entry points.
//   full = true : model map (x omega, x exposure x correction)
//   full = false: raw cubature (cubature_convolve_map_psf)
static int model_map_device(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph, int64_t S,
                            const double* expo, const double* upsf, const double* parint,
                            int64_t E, const double* rule, int64_t N, int32_t flags, void* out,
                            bool full) {
    XTS_TRY(check_wcs(wcs));
    XTS_CHECK(S > 0 && S <= (1 << 24) && E > 0 && E <= 4096, "bad S=%lld or E=%lld",
              (long long)S, (long long)E);
    XTS_CHECK(src_sph && upsf && out, "null device pointer");
    XTS_CHECK(!full || (expo && parint), "null exposure/parint");
    XTS_CHECK((flags & ~0x7) == 0, "unknown flags 0x%x", flags);
    const int H = (int)wcs->naxis1, W = (int)wcs->naxis2;
    const bool f64 = (flags & XTS_COMPUTE_F64) != 0;
    const int sehw = (flags & XTS_LAYOUT_SEHW) ? 1 : 0;
    const int of32 = (flags & XTS_OUT_F32) ? 1 : 0;

    ImageGeom ig;
    XTS_TRY(build_geometry(ctx, wcs, src_sph, (int)S, &ig));
    SepTables st;
    XTS_TRY(sep_tables(ctx, &st));
    RulePoint* rp;
    RulePoint32* rp32;
    double rho_scale;
    XTS_TRY(upload_rule(ctx, rule, N, &rp, &rp32, &rho_scale));
    if (!rule) N = 12;

    // PSF table in the cubature precision
    const int EC = f64 ? 16 : 40;
    const int Ep = (int)((E + 3) / 4 * 4);
    void* lut;
    const size_t lut_elems = (size_t)S * kND * Ep;
    XTS_TRY(slot_reserve(ctx, kSlotLut, lut_elems * (f64 ? 8 : 4), &lut));
    lut_convert_kernel<<<std::min<unsigned>(div_up(lut_elems, 256), 8192), 256, 0, ctx->stream>>>(
        upsf, (int)S, (int)E, Ep, f64 ? nullptr : (float*)lut, f64 ? (double*)lut : nullptr);

    double* mapint = nullptr;
    double* fac = nullptr;
    if (full) {
        void* acc;
        XTS_TRY(slot_reserve(ctx, kSlotAcc, (size_t)S * E * 8 * 2, &acc));
        mapint = (double*)acc;
        fac = mapint + S * E;
        zero_kernel<<<std::min<unsigned>(div_up(S * E, 256), 1024), 256, 0, ctx->stream>>>(
            mapint, (size_t)S * E);
    }

    CubArgs ca{};
    ca.corners = ig.corners;
    ca.centres = ig.centres;
    ca.omega = full ? ig.omega : nullptr;
    ca.geo = ig.geo;
    ca.lut = lut;
    ca.rule = rp;
    ca.rule32 = rp32;
    ca.inv_sep = st.inv32;
    ca.inv_sep64 = st.inv64;
    ca.mapint = mapint;
    ca.out = out;
    ca.S = (int)S; ca.H = H; ca.W = W; ca.E = (int)E; ca.Ep = Ep; ca.N = (int)N;
    ca.tiles_w = (W + kTileW - 1) / kTileW;
    ca.tiles_h = (H + kTileH - 1) / kTileH;
    const int ntiles = ca.tiles_w * ca.tiles_h;
    // About eight blocks per CU (two resident at a time), each amortising its
    // LUT staging over several tiles of the same source.
    const int nchunks = (int)((E + EC - 1) / EC);
    const int64_t work = (int64_t)ntiles * S * nchunks;
    int64_t tpb = (work + (int64_t)ctx->cu_count * 8 - 1) / ((int64_t)ctx->cu_count * 8);
    tpb = std::max<int64_t>(tpb, std::min<int64_t>(4, ntiles));
    ca.tiles_per_block = (int)std::min<int64_t>(std::max<int64_t>(tpb, 1), ntiles);
    ca.layout_sehw = sehw;
    ca.out_f32 = of32;
    ca.rho_scale = rho_scale;
    dim3 grid((ntiles + ca.tiles_per_block - 1) / ca.tiles_per_block, (unsigned)S, nchunks);
    TimedRegion tr_cub(ctx, "cubature");
    if (f64) {
        const size_t sm = (size_t)(kND * 16 + kThreads * kWinStride + kNDPad) * 8 + (size_t)N * 40;
        XTS_HIP(hipFuncSetAttribute((const void*)cubature_kernel<double, 16>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        cubature_kernel<double, 16><<<grid, kThreads, sm, ctx->stream>>>(ca);
    } else {
        const size_t sm = (size_t)(kND * 40 + kThreads * kWinStride + kNDPad) * 4 + (size_t)N * 20;
        XTS_HIP(hipFuncSetAttribute((const void*)cubature_kernel<float, 40>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        cubature_kernel<float, 40><<<grid, kThreads, sm, ctx->stream>>>(ca);
    }
    XTS_HIP(hipGetLastError());
    tr_cub.stop();

    if (full) {
        factor_kernel<<<div_up(S * E, 256), 256, 0, ctx->stream>>>(
            (int)S, (int)E, ig.geo, mapint, nullptr, upsf, parint, st.sep, expo, fac, nullptr);
        const size_t plane = (size_t)H * W;
        const size_t n = (size_t)S * E * plane;
        const unsigned blocks = std::min<unsigned>(div_up((int64_t)n, 256), 65536);
        TimedRegion tr_scale(ctx, "scale");
        if (sehw && S * E <= 65535) {
            const unsigned bx = std::max(1u, std::min<unsigned>(div_up((int64_t)plane, 1024), 64));
            if (of32)
                scale_planes_kernel<float><<<dim3(bx, (unsigned)(S * E)), 256, 0, ctx->stream>>>(
                    (float*)out, plane, expo, fac);
            else
                scale_planes_kernel<double><<<dim3(bx, (unsigned)(S * E)), 256, 0, ctx->stream>>>(
                    (double*)out, plane, expo, fac);
        } else if (of32) {
            scale_kernel<float><<<blocks, 256, 0, ctx->stream>>>((float*)out, (int)S, (int)E, plane,
                                                                 sehw, expo, fac);
        } else {
            scale_kernel<double><<<blocks, 256, 0, ctx->stream>>>((double*)out, (int)S, (int)E,
                                                                  plane, sehw, expo, fac);
        }
        XTS_HIP(hipGetLastError());
        tr_scale.stop();
    }
    return XTS_OK;
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
    SepTables st;
    XTS_TRY(sep_tables(ctx, &st));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, S * E * 8, &dout));
    factor_kernel<<<div_up(S * E, 256), 256, 0, ctx->stream>>>(
        (int)S, (int)E, (SrcGeom*)dg, nullptr, (double*)dinv, (double*)dup, (double*)dpar, st.sep,
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
