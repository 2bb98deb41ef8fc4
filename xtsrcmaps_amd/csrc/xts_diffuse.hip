// xts_diffuse.hip -- the diffuse-source model map on the GPU: a sky cube
// (the galactic diffuse MapCube) cropped to the model map's box and resampled
// onto a finer grid, convolved per energy plane with the PSF kernel through
// FFTs, sampled at the model map's pixel centres and scaled by pixel solid
// angle and exposure.
//
// Reference steps (model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97,
// a function the reference does not compile):
//   crop      fits/fits.read_allsky_cropped.cpp:150-216
//   rebin     sky_geom/sky_geom.hxx:350-380 (rebin_wcs_config, rebin_skygeom)
//   resample  math/trilerp.hpp:9-56 (trilerpEHW)
//   kernel    psf/psf.make_kernel.cpp:8-90
//   convolve  model_map/mm_fft/fft.convolve_map_psf.cpp:8-39,
//             math/convolve/fft/fft.fftw_convolve_CHW.cpp:7-90
//   scale     model_map.diffuse_src_model_map_wcs.cxx:76, 88 (solid angle,
//             exposure)
// Where those sources are unfinished or inconsistent the steps here are:
//   * the crop box bounds the cube coordinates of the whole border of the
//     model map (the reference takes two corners and never wraps, since its
//     first pixel is a min and its last a max), unwrapped across the
//     longitude seam of an all-sky CAR cube, widened by a margin;
//   * the resampled grid is the crop's corner-aligned trilinear grid
//     (trilerp.hpp's index map) and its WCS is derived from it, so the
//     kernel and the final sampling see the same grid;
//   * the kernel is centred on its middle pixel at the model map's reference
//     direction (the source position of a diffuse source,
//     source/source.collect_source_model.cpp:15) on that grid, and blends the
//     two table nodes as the point-source pixel integral does
//     (mm_cubature/convolve_pixel_psf_logsep.hpp:45-55, 72-73);
//   * the convolution is per energy plane (the reference's 3-D transform also
//     runs along the energy axis) and divides by the transform size;
//   * the convolved grid is sampled bilinearly at the model map's pixel
//     centres (the reference leaves its downsampling step empty).
//
// Kernels:
//   psf_sep_kernel       per kernel pixel: table node and blend weight
//   fft_rows_fwd_kernel  two real rows per complex transform along the rows
//                        (the resampled cube or the PSF kernel generated on
//                        the fly), half spectra out
//   fft_cols_kernel      columns: forward; the PSF's spectrum stored by
//                        column, the cube's multiplied by it and transformed
//                        back in the same launch
//   fft_rows_inv_kernel  two rows per inverse transform, real rows out
//   diffuse_sample_kernel  bilinear sampling, solid angle, exposure
// The transforms are mixed-radix (2, 3, 4, 5) Stockham passes in LDS, in place,
// FP64.
#include "xts_internal.hpp"

#include <algorithm>
#include <cstring>
#include <limits>

namespace xts {

constexpr int kFftMax = 4096;     // longest transform, elements per block
constexpr int kFftThreads = 256;

struct FftPlan {
    int n = 0, nst = 0;
    int radix[16] = {};
    int ns[16] = {};                       // product of the radices before pass s
    int tws[16] = {};                      // n / (ns R): twiddle table stride
    unsigned long long inv_t[16] = {};     // ceil(2^32 / (n / R)) of pass s
    unsigned long long inv_ns[16] = {};    // ceil(2^32 / ns) of pass s
    const double2* tw = nullptr;           // device [n]: (cos, sin)(2 pi m / n)
};

static unsigned long long inv32(unsigned d) { return ((1ULL << 32) + d - 1) / d; }

// Transform length n as radix-4, -2, -3 and -5 passes; false if n has
// another prime factor.
static bool fft_plan(int n, FftPlan* p) {
    p->n = n;
    p->nst = 0;
    int m = n;
    while (m % 4 == 0) { p->radix[p->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p->radix[p->nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p->radix[p->nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p->radix[p->nst++] = 5; m /= 5; }
    int ns = 1;
    for (int k = 0; k < p->nst; ++k) {
        const int R = p->radix[k];
        p->ns[k] = ns;
        p->tws[k] = n / (ns * R);
        p->inv_t[k] = inv32((unsigned)(n / R));
        p->inv_ns[k] = inv32((unsigned)ns);
        ns *= R;
    }
    return m == 1 && n >= 1;
}

// The smallest 2^a 3^b 5^c >= need (0 above kFftMax).
static int fft_size(long need) {
    for (long n = std::max(need, 1L); n <= kFftMax; ++n) {
        long m = n;
        for (int f : {2, 3, 5})
            while (m % f == 0) m /= f;
        if (m == 1) return (int)n;
    }
    return 0;
}

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// sg * i * a
__device__ inline double2 crot(double2 a, double sg) { return make_double2(-sg * a.y, sg * a.x); }

// R-point DFTs, y_k = sum_r v_r exp(sg 2 pi i r k / R)
template <int R> __device__ inline void dft(double2* v, double sg);

template <> __device__ inline void dft<2>(double2* v, double) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <> __device__ inline void dft<4>(double2* v, double sg) {
    const double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const double2 s13 = cadd(v[1], v[3]), d13 = crot(csub(v[1], v[3]), sg);
    v[0] = cadd(s02, s13);
    v[1] = cadd(d02, d13);
    v[2] = csub(s02, s13);
    v[3] = csub(d02, d13);
}

template <> __device__ inline void dft<3>(double2* v, double sg) {
    constexpr double kS3 = 0.86602540378443864676372317075293618;  // sqrt(3) / 2
    const double2 t1 = cadd(v[1], v[2]);
    const double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 d = crot(csub(v[1], v[2]), sg * kS3);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, d);
    v[2] = csub(t2, d);
}

template <> __device__ inline void dft<5>(double2* v, double sg) {
    // cos(2 pi / 5) = (sqrt 5 - 1) / 4, cos(4 pi / 5) = -(sqrt 5 + 1) / 4,
    // sin(2 pi / 5) = sqrt(10 + 2 sqrt 5) / 4, sin(4 pi / 5) = sqrt(10 - 2 sqrt 5) / 4
    constexpr double c1 = 0.30901699437494742410229341718281906;
    constexpr double c2 = -0.80901699437494742410229341718281906;
    constexpr double s1 = 0.95105651629515357211643933337938214;
    constexpr double s2 = 0.58778525229247312916870595463907277;
    const double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const double2 t1 = make_double2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const double2 t2 = make_double2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const double2 u1 = crot(make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sg);
    const double2 u2 = crot(make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sg);
    v[0] = make_double2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
}

// floor(x / d) from inv = ceil(2^32 / d): exact for x d < 2^32 (the error
// x (inv d - 2^32) / (d 2^32) stays below 1 / d), which holds for the
// butterfly and position indices here (x < 2^13, d < 2^13).
__device__ inline int fdiv(int x, unsigned long long inv) {
    return (int)(((unsigned long long)(unsigned)x * inv) >> 32);
}

// One Stockham pass of radix R over nb transforms of length n laid out one
// after the other in buf (ns: the product of the radices before this pass;
// nb n <= NMAX).  Butterfly j of a transform reads x[j + r n/R], twiddles by
// exp(sg 2 pi i r k / (ns R)) with k = j mod ns (from the table tw of the
// length: m = r k n / (ns R)), and writes its R outputs to (j / ns) ns R + k
// + r ns.  Every thread holds its butterflies' inputs in registers across
// the barrier, so one LDS buffer serves as input and output; NMAX bounds the
// butterflies a thread holds.
template <int R, int NMAX>
__device__ void fft_pass(double2* buf, int n, int nb, int ns, int tws,
                         unsigned long long inv_t, unsigned long long inv_ns, double sg,
                         const double2* __restrict__ tw) {
    constexpr int kMaxB = (NMAX / R + kFftThreads - 1) / kFftThreads;
    const int T = n / R;
    const int total = T * nb;
    double2 v[kMaxB][R];
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int q = threadIdx.x + b * kFftThreads;
        if (q < total) {
            const int t = fdiv(q, inv_t), j = q - t * T;
            const double2* x = buf + t * n + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = x[r * T];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int q = threadIdx.x + b * kFftThreads;
        if (q < total) {
            const int t = fdiv(q, inv_t), j = q - t * T;
            const int quo = fdiv(j, inv_ns);
            const int k = j - quo * ns;
            if (ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const double2 w = tw[r * k * tws];
                    v[b][r] = cmul(v[b][r], make_double2(w.x, sg * w.y));
                }
            }
            dft<R>(v[b], sg);
            double2* y = buf + t * n + quo * ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) y[r * ns] = v[b][r];
        }
    }
    __syncthreads();
}

// nb transforms of length p.n in buf, natural order in and out, in place;
// sg -1 forward, +1 inverse (unscaled).
template <int NMAX>
__device__ void fft_run(double2* buf, const FftPlan& p, int nb, double sg) {
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.radix[s], ns = p.ns[s], tws = p.tws[s];
        const unsigned long long it = p.inv_t[s], in = p.inv_ns[s];
        if (R == 4) fft_pass<4, NMAX>(buf, p.n, nb, ns, tws, it, in, sg, p.tw);
        else if (R == 2) fft_pass<2, NMAX>(buf, p.n, nb, ns, tws, it, in, sg, p.tw);
        else if (R == 3) fft_pass<3, NMAX>(buf, p.n, nb, ns, tws, it, in, sg, p.tw);
        else fft_pass<5, NMAX>(buf, p.n, nb, ns, tws, it, in, sg, p.tw);
    }
}

// Deterministic block sum (fixed reduction order).
__device__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    __syncthreads();
    return s;
}

// Twiddle table of a transform length n: tw[m] = (cos, sin)(2 pi m / n).
__global__ void __launch_bounds__(256) twiddle_kernel(double2* __restrict__ tw, int n) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        double s, c;
        sincospi(2.0 * m / (double)n, &s, &c);
        tw[m] = make_double2(c, s);
    }
}

// Table node i (the upper one) and blend weight t of a separation s in
// degrees, value = (1 - t) lut[i - 1] + t lut[i]
// (mm_cubature/convolve_pixel_psf_logsep.hpp:45-55, psf/psf.make_kernel.cpp:28-38).
__host__ __device__ inline void sep_node(double s, int* i, double* t) {
    const double j = s < 1e-4 ? 1.0 : 1.0 + kRecipStep * (kLn1e4 + log(s));
    int k = 1 + (int)j;
    k = k > 400 ? 400 : k;
    *i = k;
    *t = kRedm1 * (s * 1e4 * exp(kSepStep * (2.0 - k)) - 1.0);
}

// J2000 galactic frame: north galactic pole at RA 192.85948, Dec 27.12825,
// galactic longitude of the north celestial pole 122.93192 (degrees).
constexpr double kGpRa = 192.85948, kGpDec = 27.12825, kNcpL = 122.93192;

__host__ __device__ inline void frame_rotate(double lon, double lat, double pole_lon,
                                             double pole_lat, double node, double* olon,
                                             double* olat) {
    const double sd = sin(lat * kD2R), cd = cos(lat * kD2R);
    const double sp = sin(pole_lat * kD2R), cp = cos(pole_lat * kD2R);
    const double da = (lon - pole_lon) * kD2R;
    double sb = sd * sp + cd * cp * cos(da);
    sb = sb > 1.0 ? 1.0 : (sb < -1.0 ? -1.0 : sb);
    double l = node - atan2(cd * sin(da), sd * cp - cd * sp * cos(da)) * kR2D;
    l = fmod(l, 360.0);
    if (l < 0.0) l += 360.0;
    *olon = l;
    *olat = asin(sb) * kR2D;
}

// frame 0: none, 1: celestial -> galactic, 2: galactic -> celestial
__host__ __device__ inline void frame_convert(int frame, double lon, double lat, double* olon,
                                              double* olat) {
    if (frame == 1) frame_rotate(lon, lat, kGpRa, kGpDec, kNcpL, olon, olat);
    else if (frame == 2) frame_rotate(lon, lat, kNcpL, kGpDec, kGpRa, olon, olat);
    else { *olon = lon; *olat = lat; }
}

// ------------------------------------------------------------------ plan
struct DiffPlan {
    xts_diffuse_geom g{};
    Wcs roi{}, up{};
    int frame = 0;
    double up_period = 0.0;           // resampled pixels per turn of longitude (0: no wrap)
    double vlo1 = 0, vhi1 = 0, vlo2 = 0, vhi2 = 0;  // resampled coordinates inside the cube
    int c = 0;                        // kernel half width
    FftPlan P, Q;
};

// false where the model map's pixel coordinate lies outside its projection
static bool roi_to_cube(const Wcs& roi, const Wcs& cube, int frame, double p1, double p2,
                        double* q1, double* q2) {
    double lon, lat, l2, b2;
    if (wcs_pix2sph(roi, p1, p2, &lon, &lat) != 0) return false;
    frame_convert(frame, lon, lat, &l2, &b2);
    wcs_sph2pix(cube, l2, b2, q1, q2);
    return true;
}

static int make_plan(const xts_wcs* roi_x, const xts_wcs* cube_x, double target, double margin,
                     DiffPlan* pl) {
    XTS_CHECK(roi_x && cube_x, "null wcs");
    for (const xts_wcs* w : {roi_x, cube_x}) {
        XTS_CHECK(w->naxis1 > 0 && w->naxis2 > 0 && w->naxis1 <= 65536 && w->naxis2 <= 65536,
                  "bad image size %lld x %lld", (long long)w->naxis1, (long long)w->naxis2);
        XTS_CHECK(w->proj == XTS_PROJ_CAR || w->proj == XTS_PROJ_AIT, "unknown projection %d",
                  w->proj);
        XTS_CHECK(w->cdelt1 != 0.0 && w->cdelt2 != 0.0, "zero CDELT");
    }
    XTS_CHECK(std::isfinite(target) && target > 0.0, "target pixel size must be positive");
    XTS_CHECK(std::isfinite(margin) && margin >= 0.0 && margin <= 90.0,
              "margin must be in [0, 90] degrees");
    xts_diffuse_geom& g = pl->g;
    const Wcs roi = wcs_setup(*roi_x), cube = wcs_setup(*cube_x);
    pl->roi = roi;
    pl->frame = roi_x->is_galactic == cube_x->is_galactic ? 0 : (cube_x->is_galactic ? 1 : 2);
    const int back = pl->frame == 0 ? 0 : 3 - pl->frame;
    const long N1 = (long)cube_x->naxis1, N2 = (long)cube_x->naxis2;
    const long H = (long)roi_x->naxis1, W = (long)roi_x->naxis2;
    const double c1 = std::fabs(cube_x->cdelt1), c2 = std::fabs(cube_x->cdelt2);
    // a CAR cube's axis 1 is longitude: pixel coordinates repeat every
    // turn; an all-sky one wraps (its columns are read modulo naxis1)
    const double turn = cube_x->proj == XTS_PROJ_CAR && cube_x->crota2 == 0.0 ? 360.0 / c1 : 0.0;
    const bool wrap = turn > 0.0 && std::fabs(N1 * c1 - 360.0) <= 1e-6 * 360.0;
    // cube coordinates of the model map's border (pixel centres, half-pixel
    // steps), unwrapped next to its centre's across the longitude seam
    double x0, y0;
    XTS_CHECK(roi_to_cube(roi, cube, pl->frame, 0.5 * (H + 1), 0.5 * (W + 1), &x0, &y0),
              "the model map's centre lies outside its projection");
    double lo1 = x0, hi1 = x0, lo2 = y0, hi2 = y0;
    auto take = [&](double p1, double p2) {
        double q1, q2;
        if (!roi_to_cube(roi, cube, pl->frame, p1, p2, &q1, &q2)) return;
        if (turn > 0.0) q1 += turn * std::nearbyint((x0 - q1) / turn);
        lo1 = std::min(lo1, q1); hi1 = std::max(hi1, q1);
        lo2 = std::min(lo2, q2); hi2 = std::max(hi2, q2);
    };
    for (long k = 0; k <= 2 * (H - 1); ++k) {
        take(1.0 + 0.5 * k, 1.0);
        take(1.0 + 0.5 * k, (double)W);
    }
    for (long k = 0; k <= 2 * (W - 1); ++k) {
        take(1.0, 1.0 + 0.5 * k);
        take((double)H, 1.0 + 0.5 * k);
    }
    // a pole of a CAR cube's frame inside the model map: every longitude,
    // and latitude up to the pole's row
    bool pole = false;
    for (int k = 0; k < 2 && turn > 0.0; ++k) {
        double lon, lat, p1, p2, q1, q2;
        frame_convert(back, 0.0, k == 0 ? 90.0 : -90.0, &lon, &lat);
        wcs_sph2pix(roi, lon, lat, &p1, &p2);
        if (p1 >= 0.5 && p1 <= H + 0.5 && p2 >= 0.5 && p2 <= W + 0.5) {
            pole = true;
            wcs_sph2pix(cube, 0.0, k == 0 ? 90.0 : -90.0, &q1, &q2);
            lo2 = std::min(lo2, q2);
            hi2 = std::max(hi2, q2);
        }
    }
    const long mg1 = 1 + (long)std::ceil(margin / c1), mg2 = 1 + (long)std::ceil(margin / c2);
    long a1 = (long)std::floor(lo1) - mg1, b1 = (long)std::ceil(hi1) + mg1;
    long a2 = (long)std::floor(lo2) - mg2, b2 = (long)std::ceil(hi2) + mg2;
    if (pole) {
        a1 = wrap ? (long)std::floor(x0) - N1 / 2 : 1;
        b1 = wrap ? a1 + N1 - 1 : N1;
    }
    if (wrap) {
        if (b1 - a1 + 1 > N1) b1 = a1 + N1 - 1;
    } else {
        a1 = std::max(a1, 1L);
        b1 = std::min(b1, N1);
    }
    a2 = std::max(a2, 1L);
    b2 = std::min(b2, N2);
    XTS_CHECK(b1 - a1 + 1 >= 2 && b2 - a2 + 1 >= 2,
              "the model map lies outside the cube (crop %ld..%ld x %ld..%ld)", a1, b1, a2, b2);
    g.crop_a1 = a1;
    g.crop_a2 = a2;
    g.crop_n1 = b1 - a1 + 1;
    g.crop_n2 = b2 - a2 + 1;
    // rebin_wcs_config: one pixel size for both axes, the target when it is
    // finer than the cube's coarser axis, else the first axis' step
    const double pixsz = target < std::max(c1, c2) ? target : c1;
    g.m1 = std::lround(g.crop_n1 * (c1 / pixsz));
    g.m2 = std::lround(g.crop_n2 * (c2 / pixsz));
    XTS_CHECK(g.m1 >= 2 && g.m2 >= 2 && g.m1 <= kFftMax && g.m2 <= kFftMax,
              "resampled grid %lld x %lld outside 2 .. %d", (long long)g.m1, (long long)g.m2,
              kFftMax);
    g.step1 = (double)(g.crop_n1 - 1) / (double)(g.m1 - 1);
    g.step2 = (double)(g.crop_n2 - 1) / (double)(g.m2 - 1);
    // the resampled grid's WCS: its pixel u is the cube's a + (u - 1) step
    xts_wcs ux = *cube_x;
    ux.naxis1 = g.m1;
    ux.naxis2 = g.m2;
    ux.crpix1 = (cube_x->crpix1 - a1) / g.step1 + 1.0;
    ux.crpix2 = (cube_x->crpix2 - a2) / g.step2 + 1.0;
    ux.cdelt1 = cube_x->cdelt1 * g.step1;
    ux.cdelt2 = cube_x->cdelt2 * g.step2;
    g.grid = ux;
    pl->up = wcs_setup(ux);
    pl->up_period = wrap ? N1 / g.step1 : 0.0;
    const double inf = std::numeric_limits<double>::infinity();
    pl->vlo1 = wrap ? -inf : (0.5 - a1) / g.step1 + 1.0;
    pl->vhi1 = wrap ? inf : (N1 + 0.5 - a1) / g.step1 + 1.0;
    pl->vlo2 = (0.5 - a2) / g.step2 + 1.0;
    pl->vhi2 = (N2 + 0.5 - a2) / g.step2 + 1.0;
    g.wrap = wrap ? 1 : 0;
    // kernel: odd size min(m1, m2) (diffuse_src_model_map_wcs.cxx:54-61),
    // centred at the model map's reference direction
    g.nk = std::min(g.m1, g.m2);
    if (g.nk % 2 == 0) g.nk -= 1;
    pl->c = (int)((g.nk - 1) / 2);
    double lon, lat;
    frame_convert(pl->frame, roi_x->crval1, roi_x->crval2, &lon, &lat);
    wcs_sph2pix(pl->up, lon, lat, &g.kc1, &g.kc2);
    if (wrap) g.kc1 += pl->up_period * std::nearbyint((0.5 * (g.m1 + 1) - g.kc1) / pl->up_period);
    // transform sizes: rows c .. c + m - 1 of the circular convolution are
    // the linear convolution's when P >= m + c
    g.fft_p = fft_size(g.m1 + pl->c);
    g.fft_q = fft_size(g.m2 + pl->c);
    XTS_CHECK(g.fft_p > 0 && g.fft_q > 0, "transform longer than %d", kFftMax);
    fft_plan((int)g.fft_p, &pl->P);
    fft_plan((int)g.fft_q, &pl->Q);
    return XTS_OK;
}

// ------------------------------------------------------------------ kernels
// Table node and weight of each kernel pixel, [w][h]: kernel pixel (h, w) is the
// grid's pixel (kc1 + h - c, kc2 + w - c), its separation taken from the
// grid direction at (kc1, kc2) (psf/psf.make_kernel.cpp:21-37).
__global__ void __launch_bounds__(256)
psf_sep_kernel(Wcs g, double kc1, double kc2, int nk, int c, int* __restrict__ node,
               double* __restrict__ wt) {
    const int n = nk * nk;
    double d0[3];
    wcs_pix2dir(g, kc1, kc2, d0);
    // stored by kernel column: k = w nk + h (the row pass runs along h)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int w = k / nk, h = k - w * nk;
        double d[3];
        int i = 0;
        double t = 0.0;
        // a kernel pixel outside the grid's projection is 0 (node 0)
        if (wcs_pix2dir(g, kc1 + (h - c), kc2 + (w - c), d) == 0)
            sep_node(dir_diff(d, d0) * kR2D, &i, &t);
        node[k] = i;
        wt[k] = t;
    }
}

// Inputs of the row transforms.
enum { kSrcCube = 0, kSrcPsf = 1, kSrcArray = 2 };

struct RowArgs {
    FftPlan q;
    int nrows, ncols;      // nonzero rows and columns of the input
    int nb;                // row pairs per block
    int qh;                // half-spectrum length
    int rstride;           // spectrum rows per column (>= nrows)
    double2* spec;         // [E][qh][rstride]
    // cube source: trilinear resample of the crop (math/trilerp.hpp:9-56)
    const float* cube;
    long long cn1, cn2;    // cube axis lengths (axis 1 fastest in memory)
    int cE, E;             // cube planes, output planes
    long long a1, a2;      // crop start (1-based cube pixels)
    int n1, n2, wrap;      // crop size, longitude wrap
    double s1, s2, se;     // trilerp index scales
    // psf source
    const int* node;
    const double* wt;
    const double* upsf;    // [401][E]
    double* rowsum;        // [E][nrows]
    // array source: [E][nrows][ncols] FP64
    const double* arr;
};

__device__ inline long long crop_col(const RowArgs& a, int k) {
    long long p = a.a1 + k;  // 1-based cube pixel along axis 1
    if (a.wrap) p = ((p - 1) % a.cn1 + a.cn1) % a.cn1 + 1;
    return p - 1;
}

// trilerp.hpp:23-54 at output (e, i, j): i along axis 1, j along axis 2
__device__ inline double cube_value(const RowArgs& a, int e, int i, int j) {
    const double ep = e * a.se;
    const int e0 = (int)ep, e1 = min(e0 + 1, a.cE - 1);
    const double fe = ep - e0, ce = 1.0 - fe;
    const double hp = i * a.s1;
    const int h0 = (int)hp, h1 = min(h0 + 1, a.n1 - 1);
    const double fh = hp - h0, ch = 1.0 - fh;
    const double wp = j * a.s2;
    const int w0 = (int)wp, w1 = min(w0 + 1, a.n2 - 1);
    const double fw = wp - w0, cw = 1.0 - fw;
    const long long x0 = crop_col(a, h0), x1 = crop_col(a, h1);
    const long long y0 = a.a2 - 1 + w0, y1 = a.a2 - 1 + w1;
    const float* P0 = a.cube + (size_t)e0 * a.cn1 * a.cn2;
    const float* P1 = a.cube + (size_t)e1 * a.cn1 * a.cn2;
    auto at = [&](const float* P, long long x, long long y) {
        return (double)P[(size_t)y * a.cn1 + x];
    };
    const double C00 = ch * at(P0, x0, y0) + fh * at(P0, x1, y0);
    const double C01 = ch * at(P1, x0, y0) + fh * at(P1, x1, y0);
    const double C10 = ch * at(P0, x0, y1) + fh * at(P0, x1, y1);
    const double C11 = ch * at(P1, x0, y1) + fh * at(P1, x1, y1);
    const double C0 = cw * C00 + fw * C10;
    const double C1 = cw * C01 + fw * C11;
    return ce * C0 + fe * C1;
}

template <int SRC>
__device__ inline double row_value(const RowArgs& a, int e, int r, int col) {
    if (r >= a.nrows || col >= a.ncols) return 0.0;
    // the cube's lines run along axis 1 (the cube's contiguous axis): line r
    // is axis-2 index r
    if (SRC == kSrcCube) return cube_value(a, e, col, r);
    if (SRC == kSrcPsf) {
        const size_t k = (size_t)r * a.ncols + col;
        const int i = a.node[k];
        const double t = a.wt[k];
        if (i == 0) return 0.0;
        return (1.0 - t) * a.upsf[(size_t)(i - 1) * a.E + e] + t * a.upsf[(size_t)i * a.E + e];
    }
    return a.arr[((size_t)e * a.nrows + r) * a.ncols + col];
}

// Rows 2p and 2p + 1 as the real and imaginary parts of one transform; the
// two half spectra X = (Z_k + conj Z_-k) / 2, Y = (Z_k - conj Z_-k) / 2i
// (k = 0 .. q/2) to spec, stored by column: spec[e][k][r], so that the
// column pass reads whole columns and this pass writes runs of 2 nb rows.
// blockIdx.y: energy.  LDS: nb q.n values.
template <int SRC, int NMAX>
__global__ void __launch_bounds__(kFftThreads) fft_rows_fwd_kernel(RowArgs a) {
    extern __shared__ double2 lds[];
    __shared__ double red[kFftThreads / 64];
    const int e = blockIdx.y, n = a.q.n;
    const int pair0 = blockIdx.x * a.nb;
    double2* buf = lds;
    for (int t = 0; t < a.nb; ++t) {
        const int r0 = 2 * (pair0 + t);
        for (int col = threadIdx.x; col < n; col += kFftThreads)
            buf[t * n + col] =
                make_double2(row_value<SRC>(a, e, r0, col), row_value<SRC>(a, e, r0 + 1, col));
    }
    __syncthreads();
    if (SRC == kSrcPsf) {
        // row sums for the kernel's normalisation (psf.make_kernel.cpp:52, 58, 63-72)
        for (int t = 0; t < a.nb; ++t) {
            double sx = 0.0, sy = 0.0;
            for (int col = threadIdx.x; col < n; col += kFftThreads) {
                sx += buf[t * n + col].x;
                sy += buf[t * n + col].y;
            }
            sx = block_sum(sx, red);
            sy = block_sum(sy, red);
            const int r0 = 2 * (pair0 + t);
            if (threadIdx.x == 0) {
                if (r0 < a.nrows) a.rowsum[(size_t)e * a.nrows + r0] = sx;
                if (r0 + 1 < a.nrows) a.rowsum[(size_t)e * a.nrows + r0 + 1] = sy;
            }
        }
    }
    fft_run<NMAX>(buf, a.q, a.nb, -1.0);
    // column k of the spectrum: rows 2 pair0 .. 2 pair0 + 2 nb - 1, X and Y
    // of each transform next to each other
    // run = 2 nb rows, a power of two dividing the block: each thread keeps
    // its row s and steps through the columns
    const int run = 2 * a.nb;
    double2* col0 = a.spec + (size_t)e * a.qh * a.rstride + 2 * pair0;
    const int s = threadIdx.x & (run - 1), t = s >> 1;
    if (2 * pair0 + s < a.nrows)
        for (int k = threadIdx.x / run; k < a.qh; k += kFftThreads / run) {
        const double2 zk = buf[t * n + k], zm = buf[t * n + (k == 0 ? 0 : n - k)];
        col0[(size_t)k * a.rstride + s] =
            (s & 1) ? make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x))
                    : make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
    }
}

enum { kColKernel = 0, kColConvolve = 1 };

struct ColArgs {
    FftPlan p;
    int rows_in, rows_out;  // nonzero input rows, rows written back (from row c on)
    int c;                  // first output row of the circular convolution
    int qh, cb;             // half-spectrum columns, columns per block
    int rstride;            // spectrum rows per column
    double2* spec;          // [E][qh][rstride]
    double2* kspec;         // [E][qh][P], the kernel's spectrum by column
    const double* rowsum;   // [E][rows_in]: the kernel's row sums (null: sum 1)
    double scale;           // 1 / (P Q)
};

// cb spectrum columns per block (contiguous in spec).  kColKernel: forward
// transforms, scaled by 1 / (kernel sum P Q), stored in kspec.
// kColConvolve: forward, times the kernel's spectrum, inverse, rows c .. c +
// rows_out - 1 written back to rows 0 .. rows_out - 1 of the same columns.
// LDS: cb p.n values.
template <int MODE, int NMAX>
__global__ void __launch_bounds__(kFftThreads) fft_cols_kernel(ColArgs a) {
    extern __shared__ double2 lds[];
    __shared__ double red[kFftThreads / 64];
    const int e = blockIdx.y, n = a.p.n, cb = a.cb;
    const int k0 = blockIdx.x * cb;
    const int ncol = min(cb, a.qh - k0);
    double2* spec = a.spec + ((size_t)e * a.qh + k0) * a.rstride;
    double2* buf = lds;
    for (int t = 0; t < cb; ++t)
        for (int r = threadIdx.x; r < n; r += kFftThreads)
            buf[t * n + r] = (t < ncol && r < a.rows_in) ? spec[(size_t)t * a.rstride + r]
                                                         : make_double2(0.0, 0.0);
    __syncthreads();
    fft_run<NMAX>(buf, a.p, cb, -1.0);
    double2* ks = a.kspec + ((size_t)e * a.qh + k0) * n;
    if (MODE == kColKernel) {
        double tot = 0.0;
        if (a.rowsum)
            for (int h = threadIdx.x; h < a.rows_in; h += kFftThreads)
                tot += a.rowsum[(size_t)e * a.rows_in + h];
        tot = a.rowsum ? block_sum(tot, red) : 1.0;
        const double s = a.scale / tot;
        for (int idx = threadIdx.x; idx < ncol * n; idx += kFftThreads) {
            const double2 v = buf[idx];
            ks[idx] = make_double2(v.x * s, v.y * s);
        }
        return;
    }
    for (int idx = threadIdx.x; idx < cb * n; idx += kFftThreads)
        buf[idx] = idx < ncol * n ? cmul(buf[idx], ks[idx]) : make_double2(0.0, 0.0);
    __syncthreads();
    fft_run<NMAX>(buf, a.p, cb, 1.0);
    for (int t = 0; t < ncol; ++t)
        for (int r = threadIdx.x; r < a.rows_out; r += kFftThreads)
            spec[(size_t)t * a.rstride + r] = buf[t * n + r + a.c];
}

struct RowInvArgs {
    FftPlan q;
    int rows, cols_out, c;  // rows, columns written (from column c on)
    int nb, qh;
    const double2* spec;    // [E][qh][rstride]
    int rstride;
    double* out;            // [E][rows][cols_out]
};

// Rows 2p, 2p + 1 from their half spectra: Z = X + iY over all k (X_-k =
// conj X_k), inverse transform, real part row 2p, imaginary part 2p + 1.
// LDS: nb q.n values.
template <int NMAX>
__global__ void __launch_bounds__(kFftThreads) fft_rows_inv_kernel(RowInvArgs a) {
    extern __shared__ double2 lds[];
    const int e = blockIdx.y, n = a.q.n;
    const int pair0 = blockIdx.x * a.nb;
    double2* buf = lds;
    const double2* col0 = a.spec + (size_t)e * a.qh * a.rstride + 2 * pair0;
    // Z_k of transform t (k < qh); reads run along t, two rows per thread
    // nb, a power of two dividing the block: each thread keeps its transform
    // t and steps through the columns
    const int t = threadIdx.x & (a.nb - 1);
    for (int k = threadIdx.x / a.nb; k < a.qh; k += kFftThreads / a.nb) {
        const int r0 = 2 * (pair0 + t);
        double2 X = make_double2(0.0, 0.0), Y = make_double2(0.0, 0.0);
        const double2* p = col0 + (size_t)k * a.rstride + 2 * t;
        if (r0 < a.rows) X = p[0];
        if (r0 + 1 < a.rows) Y = p[1];
        buf[t * n + k] = make_double2(X.x - Y.y, X.y + Y.x);
        // Z_{n-k} = conj X_k + i conj Y_k
        if (k > 0 && n - k >= a.qh) buf[t * n + n - k] = make_double2(X.x + Y.y, Y.x - X.y);
    }
    __syncthreads();
    fft_run<NMAX>(buf, a.q, a.nb, 1.0);
    for (int tt = 0; tt < a.nb; ++tt) {
        const int r0 = 2 * (pair0 + tt);
        double* o = a.out + ((size_t)e * a.rows + r0) * a.cols_out;
        for (int j = threadIdx.x; j < a.cols_out; j += kFftThreads) {
            const double2 z = buf[tt * n + j + a.c];
            if (r0 < a.rows) o[j] = z.x;
            if (r0 + 1 < a.rows) o[a.cols_out + j] = z.y;
        }
    }
}

// The convolved grid sampled bilinearly at each model-map pixel centre,
// times the pixel's solid angle and the energy's exposure
// (model_map.diffuse_src_model_map_wcs.cxx:76, 88); zero where the pixel
// centre falls outside the cube.
template <typename OT>
__global__ void __launch_bounds__(256)
diffuse_sample_kernel(Wcs roi, Wcs up, int frame, double period, double vlo1, double vhi1,
                      double vlo2, double vhi2, int m1, int m2, int H, int W, int E,
                      const double* __restrict__ conv, const double* __restrict__ omega,
                      const double* __restrict__ expo, OT* __restrict__ out) {
    const int n = H * W;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int h = k / W, w = k - h * W;
        double lon, lat, l2, b2, v1, v2;
        const bool valid = wcs_pix2sph(roi, h + 1.0, w + 1.0, &lon, &lat) == 0;
        frame_convert(frame, lon, lat, &l2, &b2);
        wcs_sph2pix(up, l2, b2, &v1, &v2);
        if (period > 0.0) v1 += period * rint((0.5 * (m1 + 1) - v1) / period);
        const bool inside = valid && v1 >= vlo1 && v1 <= vhi1 && v2 >= vlo2 && v2 <= vhi2;
        const double x = fmin(fmax(v1 - 1.0, 0.0), (double)(m1 - 1));
        const double y = fmin(fmax(v2 - 1.0, 0.0), (double)(m2 - 1));
        const int i0 = min((int)x, m1 - 2), j0 = min((int)y, m2 - 2);
        const double fx = x - i0, fy = y - j0;
        const double w00 = (1.0 - fx) * (1.0 - fy), w01 = (1.0 - fx) * fy;
        const double w10 = fx * (1.0 - fy), w11 = fx * fy;
        const double om = omega[k];
        for (int e = 0; e < E; ++e) {
            // conv [E][m2][m1]: lines of constant axis-2 index
            const double* c = conv + ((size_t)e * m2 + j0) * m1 + i0;
            const double v = w00 * c[0] + w01 * c[m1] + w10 * c[1] + w11 * c[m1 + 1];
            out[(size_t)e * n + k] = inside ? (OT)(v * om * expo[e]) : (OT)0;
        }
    }
}

// Normalised kernel values (psf/psf.make_kernel.cpp:45-87), one block per
// energy.
__global__ void __launch_bounds__(256)
psf_kernel_values_kernel(int nk, int E, const int* __restrict__ node,
                         const double* __restrict__ wt, const double* __restrict__ upsf,
                         double* __restrict__ out) {
    __shared__ double red[4];
    const int e = blockIdx.x, n = nk * nk;
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = node[k];
        const double t = wt[k];
        const double v =
            i == 0 ? 0.0 : (1.0 - t) * upsf[(size_t)(i - 1) * E + e] + t * upsf[(size_t)i * E + e];
        out[(size_t)e * n + k] = v;
        s += v;
    }
    // the block sum above adds values in a fixed order; the row sums of
    // the FFT path sum rows first, so the two normalisations may differ in
    // the last bit
    s = block_sum(s, red);
    const double inv = 1.0 / s;
    for (int k = threadIdx.x; k < n; k += blockDim.x) out[(size_t)e * n + k] *= inv;
}

// [E][w][h] -> [E][h][w]
__global__ void __launch_bounds__(256)
kernel_transpose_kernel(int nk, int E, const double* __restrict__ in, double* __restrict__ out) {
    const size_t n = (size_t)nk * nk;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n * E;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t e = k / n, r = k - e * n;
        const int w = (int)(r / nk), h = (int)(r - (size_t)w * nk);
        out[e * n + (size_t)h * nk + w] = in[k];
    }
}

// ------------------------------------------------------------------ host
// Transforms per block for length n: at most 1536 values a block (LDS
// 24 KB), so that several blocks share a CU; NMAX: the kernel variant whose
// register arrays hold a block of nb n values.
static int choose_nb(int n) {
    int nb = 1;
    while (nb < 8 && 2 * nb * n <= 1536) nb *= 2;
    return nb;
}
static int pick_nmax(int elems) {
    return elems <= 1024 ? 1024 : (elems <= 1536 ? 1536 : (elems <= 2048 ? 2048 : 4096));
}
template <typename F>
static F by_nmax(int elems, F f1024, F f1536, F f2048, F f4096) {
    const int m = pick_nmax(elems);
    return m == 1024 ? f1024 : (m == 1536 ? f1536 : (m == 2048 ? f2048 : f4096));
}
template <int SRC>
static void (*rows_fwd_fn(int elems))(RowArgs) {
    return by_nmax(elems, fft_rows_fwd_kernel<SRC, 1024>, fft_rows_fwd_kernel<SRC, 1536>,
                   fft_rows_fwd_kernel<SRC, 2048>, fft_rows_fwd_kernel<SRC, 4096>);
}
template <int MODE>
static void (*cols_fn(int elems))(ColArgs) {
    return by_nmax(elems, fft_cols_kernel<MODE, 1024>, fft_cols_kernel<MODE, 1536>,
                   fft_cols_kernel<MODE, 2048>, fft_cols_kernel<MODE, 4096>);
}
static void (*rows_inv_fn(int elems))(RowInvArgs) {
    return by_nmax(elems, fft_rows_inv_kernel<1024>, fft_rows_inv_kernel<1536>,
                   fft_rows_inv_kernel<2048>, fft_rows_inv_kernel<4096>);
}
static int spec_stride(int rows) { return (rows + 7) & ~7; }

// The twiddle table of p into tw (p.n entries).
static int make_twiddles(xts_ctx* ctx, FftPlan* p, double2* tw) {
    twiddle_kernel<<<div_up(p->n, 256), 256, 0, ctx->stream>>>(tw, p->n);
    XTS_HIP(hipGetLastError());
    p->tw = tw;
    return XTS_OK;
}

static size_t lds_bytes(int nb, int n) { return (size_t)nb * n * sizeof(double2); }

template <typename A>
static int launch_lds(void (*kernel)(A), dim3 grid, size_t lds, hipStream_t st, A args) {
    XTS_TRY(allow_lds((const void*)kernel, lds));
    void* kargs[] = {&args};
    XTS_HIP(hipLaunchKernel((const void*)kernel, grid, dim3(kFftThreads), kargs, lds, st));
    return XTS_OK;
}

// Row pass of E planes (rows ra.nrows) into spec, stored by column with
// ra.rstride rows a column.
static int rows_forward(hipStream_t st, RowArgs ra, const FftPlan& Qp, int E) {
    const int Q = Qp.n;
    ra.q = Qp;
    ra.nb = choose_nb(Q);
    ra.qh = Q / 2 + 1;
    ra.E = E;
    const dim3 g(div_up(div_up(ra.nrows, 2), ra.nb), E);
    const size_t lds = lds_bytes(ra.nb, Q);
    const int el = ra.nb * Q;
    if (ra.node) return launch_lds(rows_fwd_fn<kSrcPsf>(el), g, lds, st, ra);
    if (ra.arr) return launch_lds(rows_fwd_fn<kSrcArray>(el), g, lds, st, ra);
    return launch_lds(rows_fwd_fn<kSrcCube>(el), g, lds, st, ra);
}

// Column pass computing the kernel's spectrum kspec [E][qh][P] from its row
// spectra in spec (rows, stride rstride); rowsum: its row sums, or null.
static int kernel_spectrum(hipStream_t st, const FftPlan& Pp, int qh, int E, int rows,
                           int rstride, double2* spec, const double* rowsum, double2* kspec,
                           double scale) {
    ColArgs ca{};
    ca.p = Pp;
    ca.rows_in = rows;
    ca.rows_out = 0;
    ca.qh = qh;
    ca.cb = choose_nb(Pp.n);
    ca.rstride = rstride;
    ca.spec = spec;
    ca.kspec = kspec;
    ca.rowsum = rowsum;
    ca.scale = scale;
    return launch_lds(cols_fn<kColKernel>(ca.cb * Pp.n), dim3(div_up(qh, ca.cb), E),
                      lds_bytes(ca.cb, Pp.n), st, ca);
}

// The kernel's spectrum into kspec [E][P/2 + 1][Q], on stream st: lines
// along axis 1 (kernel columns w, transforms of length P), then across them
// (length Q); spec is scratch of at least E x (P/2 + 1) x spec_stride(nk)
// values.
static int psf_spectrum(hipStream_t st, const DiffPlan& pl, const double* upsf, int E,
                        int* node, double* wt, double* rowsum, double2* spec, double2* kspec) {
    const xts_diffuse_geom& g = pl.g;
    const int nk = (int)g.nk, P = (int)g.fft_p, Q = (int)g.fft_q;
    psf_sep_kernel<<<std::min<unsigned>(div_up((int64_t)nk * nk, 256), 4096), 256, 0, st>>>(
        pl.up, g.kc1, g.kc2, nk, pl.c, node, wt);
    XTS_HIP(hipGetLastError());
    RowArgs ra{};
    ra.nrows = nk;
    ra.ncols = nk;
    ra.rstride = spec_stride(nk);
    ra.spec = spec;
    ra.node = node;
    ra.wt = wt;
    ra.upsf = upsf;
    ra.rowsum = rowsum;
    XTS_TRY(rows_forward(st, ra, pl.P, E));
    return kernel_spectrum(st, pl.Q, P / 2 + 1, E, nk, ra.rstride, spec, rowsum, kspec,
                           1.0 / ((double)P * Q));
}

// Buffers of one diffuse map in the context's slots.
struct DiffBufs {
    int* node;
    double* wt;
    double* rowsum;
    double2* tw;          // twiddles of P, then of Q
    double2* spec;        // row spectra (the kernel's, then the image's)
    double2* kspec;       // the kernel's spectrum
    double* conv;
};

static int diffuse_bufs(xts_ctx* ctx, const DiffPlan& pl, int E, DiffBufs* b) {
    const xts_diffuse_geom& g = pl.g;
    const size_t nk = (size_t)g.nk, P = (size_t)g.fft_p, Q = (size_t)g.fft_q;
    const size_t qh = P / 2 + 1;
    const size_t rows = (size_t)spec_stride((int)std::max<int64_t>(g.m2, g.nk));
    void *p0, *p1, *p2, *p3;
    const size_t sep_bytes = arena_bytes(nk * nk * 4) + arena_bytes(nk * nk * 8) +
                             arena_bytes((size_t)E * nk * 8) + arena_bytes((P + Q) * 16);
    XTS_TRY(slot_reserve(ctx, kSlotDiffSep, sep_bytes, &p0));
    XTS_TRY(slot_reserve(ctx, kSlotDiffSpec, (size_t)E * qh * rows * 16, &p1));
    XTS_TRY(slot_reserve(ctx, kSlotDiffKspec, (size_t)E * qh * Q * 16, &p2));
    XTS_TRY(slot_reserve(ctx, kSlotDiffConv, (size_t)E * g.m1 * g.m2 * 8, &p3));
    char* c = (char*)p0;
    b->node = (int*)c;
    c += arena_bytes(nk * nk * 4);
    b->wt = (double*)c;
    c += arena_bytes(nk * nk * 8);
    b->rowsum = (double*)c;
    c += arena_bytes((size_t)E * nk * 8);
    b->tw = (double2*)c;
    b->spec = (double2*)p1;
    b->kspec = (double2*)p2;
    b->conv = (double*)p3;
    return XTS_OK;
}

// Convolution of E planes (the cube's resample, or an array: ra.arr) with
// the kernel whose spectrum is in kspec: rows c .. c + rows_out - 1 and
// columns c .. c + cols_out - 1 of the linear convolution into out
// [E][rows_out][cols_out].  spec holds E x qh x spec_stride(max(rows,
// rows_out)) values.
static int convolve(xts_ctx* ctx, const FftPlan& Pp, const FftPlan& Qp, RowArgs ra, int E,
                    int rows_out, int cols_out, int c, double2* spec, const double2* kspec,
                    double* out) {
    const int P = Pp.n, Q = Qp.n;
    const int qh = Q / 2 + 1;
    ra.rstride = spec_stride(std::max(ra.nrows, rows_out));
    ra.spec = spec;
    XTS_TRY(rows_forward(ctx->stream, ra, Qp, E));
    ColArgs ca{};
    ca.p = Pp;
    ca.rows_in = ra.nrows;
    ca.rows_out = rows_out;
    ca.c = c;
    ca.qh = qh;
    ca.cb = choose_nb(P);
    ca.rstride = ra.rstride;
    ca.spec = spec;
    ca.kspec = const_cast<double2*>(kspec);
    {
        TimedRegion tr(ctx, "diffuse_cols");
        XTS_TRY(launch_lds(cols_fn<kColConvolve>(ca.cb * P), dim3(div_up(qh, ca.cb), E),
                           lds_bytes(ca.cb, P), ctx->stream, ca));
    }
    RowInvArgs ri{};
    ri.q = Qp;
    ri.rows = rows_out;
    ri.cols_out = cols_out;
    ri.c = c;
    ri.nb = choose_nb(Q);
    ri.qh = qh;
    ri.spec = spec;
    ri.rstride = ra.rstride;
    ri.out = out;
    return launch_lds(rows_inv_fn(ri.nb * Q), dim3(div_up(div_up(rows_out, 2), ri.nb), E),
                      lds_bytes(ri.nb, Q), ctx->stream, ri);
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_diffuse_geometry(const xts_wcs* roi, const xts_wcs* cube, double target_pix_size,
                         double margin_deg, xts_diffuse_geom* out) {
    XTS_CHECK(out != nullptr, "null out");
    DiffPlan pl;
    XTS_TRY(make_plan(roi, cube, target_pix_size, margin_deg, &pl));
    *out = pl.g;
    return XTS_OK;
}

int xts_diffuse_src_model_map_wcs_dev(xts_ctx* ctx, const xts_wcs* roi, const xts_wcs* cube_wcs,
                                      const float* cube, int64_t n_planes, const double* upsf,
                                      const double* expo, int64_t E, double target_pix_size,
                                      double margin_deg, void* out, int32_t flags) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(cube && upsf && expo && out, "null device pointer");
    XTS_CHECK(n_planes >= 1 && n_planes <= 65536 && E >= 1 && E <= 4096,
              "bad plane counts (cube %lld, energies %lld)", (long long)n_planes, (long long)E);
    XTS_CHECK((flags & ~XTS_OUT_F32) == 0, "unknown flags 0x%x", flags);
    DiffPlan pl;
    XTS_TRY(make_plan(roi, cube_wcs, target_pix_size, margin_deg, &pl));
    const xts_diffuse_geom& g = pl.g;
    const int H = (int)roi->naxis1, W = (int)roi->naxis2;
    XTS_CHECK((int64_t)H * W <= (int64_t)1 << 30, "image too large");
    const double* omega;
    XTS_TRY(image_solid_angle(ctx, roi, &omega));
    DiffBufs b;
    XTS_TRY(diffuse_bufs(ctx, pl, (int)E, &b));
    XTS_TRY(make_twiddles(ctx, &pl.P, b.tw));
    XTS_TRY(make_twiddles(ctx, &pl.Q, b.tw + pl.P.n));
    {
        TimedRegion tr(ctx, "diffuse_kernel");
        XTS_TRY(psf_spectrum(ctx->stream, pl, upsf, (int)E, b.node, b.wt, b.rowsum, b.spec,
                             b.kspec));
    }
    {
        TimedRegion tr(ctx, "diffuse_fft");
        // lines of constant axis-2 index, along axis 1
        RowArgs ra{};
        ra.nrows = (int)g.m2;
        ra.ncols = (int)g.m1;
        ra.cube = cube;
        ra.cn1 = cube_wcs->naxis1;
        ra.cn2 = cube_wcs->naxis2;
        ra.cE = (int)n_planes;
        ra.a1 = g.crop_a1;
        ra.a2 = g.crop_a2;
        ra.n1 = (int)g.crop_n1;
        ra.n2 = (int)g.crop_n2;
        ra.wrap = (int)g.wrap;
        // trilerp.hpp:23-25: (src - 1) / (dst - 1) per axis
        ra.s1 = (double)(g.crop_n1 - 1) / (double)(g.m1 - 1);
        ra.s2 = (double)(g.crop_n2 - 1) / (double)(g.m2 - 1);
        ra.se = E > 1 ? (double)(n_planes - 1) / (double)(E - 1) : 0.0;
        XTS_TRY(convolve(ctx, pl.Q, pl.P, ra, (int)E, (int)g.m2, (int)g.m1, pl.c, b.spec,
                         b.kspec, b.conv));
    }
    {
        TimedRegion tr(ctx, "diffuse_sample");
        const unsigned grid = std::min<unsigned>(div_up((int64_t)H * W, 256), 8192);
        if (flags & XTS_OUT_F32)
            diffuse_sample_kernel<float><<<grid, 256, 0, ctx->stream>>>(
                pl.roi, pl.up, pl.frame, pl.up_period, pl.vlo1, pl.vhi1, pl.vlo2, pl.vhi2,
                (int)g.m1, (int)g.m2, H, W, (int)E, b.conv, omega, expo, (float*)out);
        else
            diffuse_sample_kernel<double><<<grid, 256, 0, ctx->stream>>>(
                pl.roi, pl.up, pl.frame, pl.up_period, pl.vlo1, pl.vhi1, pl.vlo2, pl.vhi2,
                (int)g.m1, (int)g.m2, H, W, (int)E, b.conv, omega, expo, (double*)out);
        XTS_HIP(hipGetLastError());
    }
    return XTS_OK;
}

int xts_diffuse_src_model_map_wcs(xts_ctx* ctx, const xts_wcs* roi, const xts_wcs* cube_wcs,
                                  const float* cube, int64_t n_planes, const double* upsf,
                                  const double* expo, int64_t E, double target_pix_size,
                                  double margin_deg, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(cube && upsf && expo && out && cube_wcs && roi, "null host pointer");
    XTS_CHECK(n_planes >= 1 && n_planes <= 65536 && E >= 1 && E <= 4096, "bad plane counts");
    const size_t cb = (size_t)n_planes * cube_wcs->naxis1 * cube_wcs->naxis2 * 4;
    const size_t ob = (size_t)E * roi->naxis1 * roi->naxis2 * 8;
    void *dc, *du, *de, *dout;
    XTS_TRY(to_device(ctx, kSlotIn0, cube, cb, &dc));
    XTS_TRY(to_device(ctx, kSlotIn1, upsf, (size_t)kND * E * 8, &du));
    XTS_TRY(to_device(ctx, kSlotIn2, expo, (size_t)E * 8, &de));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, ob, &dout));
    XTS_TRY(xts_diffuse_src_model_map_wcs_dev(ctx, roi, cube_wcs, (const float*)dc, n_planes,
                                              (const double*)du, (const double*)de, E,
                                              target_pix_size, margin_deg, dout, 0));
    XTS_HIP(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_psf_make_kernel(xts_ctx* ctx, int64_t nk, const xts_wcs* grid, double kc1, double kc2,
                        const double* upsf, int64_t E, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(grid && upsf && out, "null pointer");
    XTS_CHECK(nk >= 1 && nk % 2 == 1 && nk <= kFftMax, "kernel size %lld must be odd, 1 .. %d",
              (long long)nk, kFftMax);
    XTS_CHECK(E >= 1 && E <= 4096, "bad energy count %lld", (long long)E);
    XTS_CHECK(grid->proj == XTS_PROJ_CAR || grid->proj == XTS_PROJ_AIT, "unknown projection");
    const size_t n = (size_t)nk * nk;
    void *du, *dsep, *dout, *dtmp;
    XTS_TRY(to_device(ctx, kSlotIn0, upsf, (size_t)kND * E * 8, &du));
    XTS_TRY(slot_reserve(ctx, kSlotDiffSep, arena_bytes(n * 4) + arena_bytes(n * 8), &dsep));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, (size_t)E * n * 8, &dout));
    XTS_TRY(slot_reserve(ctx, kSlotOut1, (size_t)E * n * 8, &dtmp));
    int* node = (int*)dsep;
    double* wt = (double*)((char*)dsep + arena_bytes(n * 4));
    psf_sep_kernel<<<std::min<unsigned>(div_up((int64_t)n, 256), 4096), 256, 0, ctx->stream>>>(
        wcs_setup(*grid), kc1, kc2, (int)nk, (int)((nk - 1) / 2), node, wt);
    psf_kernel_values_kernel<<<(unsigned)E, 256, 0, ctx->stream>>>(
        (int)nk, (int)E, node, wt, (const double*)du, (double*)dtmp);
    kernel_transpose_kernel<<<std::min<unsigned>(div_up((int64_t)E * n, 256), 8192), 256, 0,
                              ctx->stream>>>((int)nk, (int)E, (const double*)dtmp, (double*)dout);
    XTS_HIP(hipGetLastError());
    XTS_HIP(hipMemcpyAsync(out, dout, (size_t)E * n * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_fft_convolve_map_psf(xts_ctx* ctx, const double* image, int64_t C, int64_t H, int64_t W,
                             const double* kernel, int64_t KH, int64_t KW, double* out) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(image && kernel && out, "null pointer");
    XTS_CHECK(C >= 1 && C <= 4096 && H >= 1 && W >= 1 && KH >= 1 && KW >= 1,
              "bad sizes C=%lld H=%lld W=%lld KH=%lld KW=%lld", (long long)C, (long long)H,
              (long long)W, (long long)KH, (long long)KW);
    const int64_t OH = H + KH - 1, OW = W + KW - 1;
    const int P = fft_size(OH), Q = fft_size(OW);
    XTS_CHECK(P > 0 && Q > 0, "output %lld x %lld needs a transform longer than %d",
              (long long)OH, (long long)OW, kFftMax);
    FftPlan Pp, Qp;
    fft_plan(P, &Pp);
    fft_plan(Q, &Qp);
    const int qh = Q / 2 + 1;
    const int64_t rows = spec_stride((int)std::max(OH, std::max(H, KH)));
    void *dimg, *dker, *dsp, *dks, *dout, *dtw;
    XTS_TRY(slot_reserve(ctx, kSlotDiffSep, (size_t)(P + Q) * 16, &dtw));
    XTS_TRY(make_twiddles(ctx, &Pp, (double2*)dtw));
    XTS_TRY(make_twiddles(ctx, &Qp, (double2*)dtw + P));
    XTS_TRY(to_device(ctx, kSlotIn0, image, (size_t)C * H * W * 8, &dimg));
    XTS_TRY(to_device(ctx, kSlotIn1, kernel, (size_t)C * KH * KW * 8, &dker));
    XTS_TRY(slot_reserve(ctx, kSlotDiffSpec, (size_t)C * qh * rows * 16, &dsp));
    XTS_TRY(slot_reserve(ctx, kSlotDiffKspec, (size_t)C * qh * P * 16, &dks));
    XTS_TRY(slot_reserve(ctx, kSlotOut0, (size_t)C * OH * OW * 8, &dout));
    // the kernel's spectrum, divided by P Q only
    RowArgs ka{};
    ka.nrows = (int)KH;
    ka.ncols = (int)KW;
    ka.rstride = spec_stride((int)KH);
    ka.spec = (double2*)dsp;
    ka.arr = (const double*)dker;
    XTS_TRY(rows_forward(ctx->stream, ka, Qp, (int)C));
    XTS_TRY(kernel_spectrum(ctx->stream, Pp, qh, (int)C, (int)KH, ka.rstride, (double2*)dsp,
                            nullptr, (double2*)dks, 1.0 / ((double)P * Q)));
    RowArgs ra{};
    ra.nrows = (int)H;
    ra.ncols = (int)W;
    ra.arr = (const double*)dimg;
    XTS_TRY(convolve(ctx, Pp, Qp, ra, (int)C, (int)OH, (int)OW, 0, (double2*)dsp,
                     (const double2*)dks, (double*)dout));
    XTS_HIP(hipMemcpyAsync(out, dout, (size_t)C * OH * OW * 8, hipMemcpyDeviceToHost,
                           ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

}  // extern "C"
