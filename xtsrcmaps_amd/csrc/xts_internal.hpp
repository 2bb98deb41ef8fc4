// xts_internal.hpp -- shared pieces of the HIP implementation: context,
// error reporting, device arena, separation-grid constants and the FP64 WCS
// math used both by kernels and by host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/xtsrcmaps_amd.h"

namespace xts {

// ---------------------------------------------------------------- constants
// Separation grid (psf/psf.hxx:24-31, misc/misc.hxx:22-39): node 0 at 0 and
// node i >= 1 at 1e-4 deg * exp((i - 1) ln(7e5) / 399).
constexpr int kND = 401;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kTwoPi = 6.28318530717958647692528676655900577;
constexpr double kHalfPi = 1.57079632679489661923132169163975144;
constexpr double kD2R = 0.01745329251994329576923690768488612713;
constexpr double kR2D = 57.2957795130823208767981548141051703324;
constexpr double kSepStep = 0.033731417579011382769913057686378448039;
constexpr double kRecipStep = 29.645952402019046699874076387601854602589;
constexpr double kRedm1 = 29.148763300179999028597222671574207003288;
constexpr double kLn1e4 = 9.2103403719761827360719658187374568304044;

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define XTS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return ::xts::fail(e_ == hipErrorOutOfMemory ? XTS_ERR_ALLOC     \
                                                         : XTS_ERR_DEVICE,   \
                               "%s: %s (%s:%d)", #call,                      \
                               hipGetErrorString(e_), __FILE__, __LINE__);   \
    } while (0)

#define XTS_CHECK(cond, ...)                                                 \
    do {                                                                     \
        if (!(cond)) return ::xts::fail(XTS_ERR_INVALID, __VA_ARGS__);       \
    } while (0)

#define XTS_TRY(expr)                                                        \
    do {                                                                     \
        int rc_ = (expr);                                                    \
        if (rc_ != XTS_OK) return rc_;                                       \
    } while (0)

// ---------------------------------------------------------------- arena
// A bump allocator over one device buffer per slot.  Slots keep their
// buffer between calls and grow (free + malloc) only when a call needs more,
// so steady-state calls allocate nothing.
struct Slot {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum SlotId {
    kSlotIn0 = 0, kSlotIn1, kSlotIn2, kSlotIn3, kSlotIn4, kSlotIn5,
    kSlotIn6, kSlotIn7, kSlotOut0, kSlotOut1, kSlotOut2, kSlotOut3,
    kSlotGeom, kSlotLut, kSlotAcc, kSlotWork0, kSlotWork1, kSlotWork2,
    kSlotWork3, kSlotTab, kSlotRows, kSlotCount
};

}  // namespace xts

struct xts_timing_rec {
    const char* name;
    hipEvent_t start, stop;
};

struct xts_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    xts::Slot slots[xts::kSlotCount];
    int cu_count = 0;
    int timing = 0;                        // record events around kernels
    int tune_tiles_per_block = 0;          // 0: chosen per call (row-weight pass)
    std::vector<xts_timing_rec> timing_recs;
    std::vector<hipEvent_t> event_pool;
};

namespace xts {

int slot_reserve(xts_ctx* ctx, int slot, size_t bytes, void** out);

// Brackets kernel launches with events on the context stream when timing is
// enabled (xts_ctx_enable_timing); no effect otherwise.
hipEvent_t pooled_event(xts_ctx* ctx);
struct TimedRegion {
    xts_ctx* ctx;
    const char* name;
    hipEvent_t start = nullptr;
    bool open = false;
    TimedRegion(xts_ctx* c, const char* n) : ctx(c), name(n) {
        if (ctx->timing) {
            start = pooled_event(ctx);
            if (start && hipEventRecord(start, ctx->stream) == hipSuccess) open = true;
        }
    }
    void stop() {
        if (!open) return;
        hipEvent_t e = pooled_event(ctx);
        if (e && hipEventRecord(e, ctx->stream) == hipSuccess)
            ctx->timing_recs.push_back({name, start, e});
        open = false;
    }
    ~TimedRegion() { stop(); }
};
int to_device(xts_ctx* ctx, int slot, const void* host, size_t bytes, void** dev);
int activate(xts_ctx* ctx);

// ---------------------------------------------------------------- WCS math
// Plain-data WCS with everything the projection needs precomputed on the
// host (xts_wcs_setup).  Greisen & Calabretta 2002 (CROTA2 convention),
// Calabretta & Greisen 2002 (CAR, AIT; native reference point (0, 0);
// default LONPOLE / LATPOLE).
struct Wcs {
    double crpix1, crpix2;
    double m11, m12, m21, m22;  // pixel offset -> intermediate (deg)
    double i11, i12, i21, i22;  // inverse
    double alphap, deltap, phip;
    double sdp, cdp;            // sin/cos of deltap
    int proj;
    long nh, nw;
};

Wcs wcs_setup(const xts_wcs& w);

__host__ __device__ inline double sind(double a) { return sin(a * kD2R); }
__host__ __device__ inline double cosd(double a) { return cos(a * kD2R); }

// pixel (1-based FITS coordinates) -> (lon, lat) in degrees; lon in [0, 360)
__host__ __device__ inline int wcs_pix2sph(const Wcs& w, double p1, double p2,
                                           double* lon, double* lat) {
    const double dp1 = p1 - w.crpix1, dp2 = p2 - w.crpix2;
    const double x = w.m11 * dp1 + w.m12 * dp2;
    const double y = w.m21 * dp1 + w.m22 * dp2;
    double phi, theta;
    int rc = 0;
    if (w.proj == XTS_PROJ_CAR) {
        phi = x;
        theta = y;
        if (fabs(theta) > 90.0) rc = -1;
    } else {
        const double r0 = kR2D;
        double u = 1.0 - x * x / (16.0 * r0 * r0) - y * y / (4.0 * r0 * r0);
        if (u < 0.0) { rc = -1; u = 0.0; }
        const double z = sqrt(u);
        double s = z * y / r0;
        if (s > 1.0) { s = 1.0; rc = -1; }
        if (s < -1.0) { s = -1.0; rc = -1; }
        theta = asin(s) * kR2D;
        const double xp = 2.0 * z * z - 1.0, yp = z * x / (2.0 * r0);
        phi = (xp == 0.0 && yp == 0.0) ? 0.0 : 2.0 * atan2(yp, xp) * kR2D;
    }
    // native -> celestial (Paper II eq. 2)
    const double dphi = (phi - w.phip) * kD2R;
    const double st = sin(theta * kD2R), ct = cos(theta * kD2R);
    const double cdphi = cos(dphi);
    const double xx = st * w.cdp - ct * w.sdp * cdphi;
    const double yy = -ct * sin(dphi);
    double a = w.alphap + atan2(yy, xx) * kR2D;
    double sl = st * w.sdp + ct * w.cdp * cdphi;
    sl = sl > 1.0 ? 1.0 : (sl < -1.0 ? -1.0 : sl);
    while (a < 0.0) a += 360.0;
    while (a >= 360.0) a -= 360.0;
    *lon = a;
    *lat = asin(sl) * kR2D;
    return rc;
}

__host__ __device__ inline void sph2dir(double lon, double lat, double* d) {
    const double ca = cos(lon * kD2R), cd = cos(lat * kD2R);
    const double sa = sin(lon * kD2R), sd = sin(lat * kD2R);
    d[0] = ca * cd;
    d[1] = sa * cd;
    d[2] = sd;
}

__host__ __device__ inline int wcs_pix2dir(const Wcs& w, double p1, double p2,
                                           double* d) {
    double lon, lat;
    const int rc = wcs_pix2sph(w, p1, p2, &lon, &lat);
    sph2dir(lon, lat, d);
    return rc;
}

__host__ __device__ inline void wcs_sph2pix(const Wcs& w, double lon,
                                            double lat, double* p1,
                                            double* p2) {
    if (lon > 180.0) lon -= 360.0;
    if (lon < -180.0) lon += 360.0;
    // celestial -> native (Paper II eq. 5)
    const double da = (lon - w.alphap) * kD2R;
    const double sd = sin(lat * kD2R), cd = cos(lat * kD2R);
    const double cda = cos(da);
    const double xx = sd * w.cdp - cd * w.sdp * cda;
    const double yy = -cd * sin(da);
    double phi = w.phip + atan2(yy, xx) * kR2D;
    while (phi > 180.0) phi -= 360.0;
    while (phi < -180.0) phi += 360.0;
    double s = sd * w.sdp + cd * w.cdp * cda;
    s = s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s);
    const double theta = asin(s) * kR2D;
    double x, y;
    if (w.proj == XTS_PROJ_CAR) {
        x = phi;
        y = theta;
    } else {
        const double r0 = kR2D;
        const double ct = cos(theta * kD2R), st = sin(theta * kD2R);
        const double g = sqrt(2.0 * r0 * r0 / (1.0 + ct * cos(0.5 * phi * kD2R)));
        x = 2.0 * g * ct * sin(0.5 * phi * kD2R);
        y = g * st;
    }
    *p1 = w.crpix1 + w.i11 * x + w.i12 * y;
    *p2 = w.crpix2 + w.i21 * x + w.i22 * y;
}

// angular distance (radians) between unit vectors: 2 asin(|a - b| / 2)
// (sky_geom/sky_geom.hxx:305-310)
__host__ __device__ inline double dir_diff(const double* a, const double* b) {
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return 2.0 * asin(0.5 * sqrt(x * x + y * y + z * z));
}

// ---------------------------------------------------------------- helpers
inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// The reference's production cubature rule: Festa-Sommariva degree-7,
// 12-point rule on [-1, 1]^2 (mm_cubature/cubature.hpp:209-225).
extern const double kSquareDeg7[36];

}  // namespace xts
