// xts_wcs.hpp -- FP64 WCS math of the model-map images, shared by the GPU
// kernels and by host code (plain C++ when compiled without HIP).
// Greisen & Calabretta 2002 (linear part, CROTA2 convention), Calabretta &
// Greisen 2002 (CAR and AIT, native reference point (0, 0), default
// LONPOLE / LATPOLE), as the reference evaluates them through wcslib
// (sky_geom/sky_geom.hxx:91-121, 179-310).
#pragma once

#include <cmath>

#include "../../include/xtsrcmaps_amd.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define XTS_HD __host__ __device__
#else
#define XTS_HD
#endif

namespace xts {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kTwoPi = 6.28318530717958647692528676655900577;
constexpr double kHalfPi = 1.57079632679489661923132169163975144;
constexpr double kD2R = 0.01745329251994329576923690768488612713;
constexpr double kR2D = 57.2957795130823208767981548141051703324;

// Plain-data WCS with everything the projection needs precomputed on the
// host (wcs_setup).  Greisen & Calabretta 2002 (CROTA2 convention),
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

XTS_HD inline double sind(double a) { return sin(a * kD2R); }
XTS_HD inline double cosd(double a) { return cos(a * kD2R); }

// pixel (1-based FITS coordinates) -> (lon, lat) in degrees; lon in [0, 360)
XTS_HD inline int wcs_pix2sph(const Wcs& w, double p1, double p2, double* lon, double* lat) {
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

XTS_HD inline void sph2dir(double lon, double lat, double* d) {
    const double ca = cos(lon * kD2R), cd = cos(lat * kD2R);
    const double sa = sin(lon * kD2R), sd = sin(lat * kD2R);
    d[0] = ca * cd;
    d[1] = sa * cd;
    d[2] = sd;
}

XTS_HD inline int wcs_pix2dir(const Wcs& w, double p1, double p2, double* d) {
    double lon, lat;
    const int rc = wcs_pix2sph(w, p1, p2, &lon, &lat);
    sph2dir(lon, lat, d);
    return rc;
}

XTS_HD inline void wcs_sph2pix(const Wcs& w, double lon, double lat, double* p1, double* p2) {
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
XTS_HD inline double dir_diff(const double* a, const double* b) {
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return 2.0 * asin(0.5 * sqrt(x * x + y * y + z * z));
}

inline Wcs wcs_setup(const xts_wcs& in) {
    Wcs w{};
    w.crpix1 = in.crpix1;
    w.crpix2 = in.crpix2;
    w.proj = in.proj;
    w.nh = (long)in.naxis1;
    w.nw = (long)in.naxis2;
    // CROTA2 -> linear transform: x = cdelt1 cos r dp1 - cdelt2 sin r dp2,
    // y = cdelt1 sin r dp1 + cdelt2 cos r dp2 (Greisen & Calabretta 2002).
    const double cr = cosd(in.crota2), sr = sind(in.crota2);
    w.m11 = in.cdelt1 * cr;
    w.m12 = -in.cdelt2 * sr;
    w.m21 = in.cdelt1 * sr;
    w.m22 = in.cdelt2 * cr;
    const double det = w.m11 * w.m22 - w.m12 * w.m21;
    w.i11 = w.m22 / det;
    w.i12 = -w.m12 / det;
    w.i21 = -w.m21 / det;
    w.i22 = w.m11 / det;
    // Celestial pole of the native frame for a reference point at native
    // (phi0, theta0) = (0, 0): default LONPOLE is 0 when delta0 >= theta0,
    // else 180; of the two pole latitudes choose the one nearest LATPOLE=90.
    const double a0 = in.crval1 > 180.0 ? in.crval1 - 360.0 : in.crval1;
    const double d0 = in.crval2;
    const double phip = d0 >= 0.0 ? 0.0 : 180.0;
    const double x = cosd(phip), y = 0.0;
    const double z = sqrt(x * x + y * y);
    const double u = atan2(y, x) * kR2D;
    const double v = acos(sind(d0) / z) * kR2D;
    double c1 = u + v, c2 = u - v;
    if (c1 > 180.0) c1 -= 360.0; else if (c1 < -180.0) c1 += 360.0;
    if (c2 > 180.0) c2 -= 360.0; else if (c2 < -180.0) c2 += 360.0;
    const bool ok1 = fabs(c1) <= 90.0 + 1e-10, ok2 = fabs(c2) <= 90.0 + 1e-10;
    double dp = (ok1 && ok2) ? (fabs(c1 - 90.0) < fabs(c2 - 90.0) ? c1 : c2) : (ok1 ? c1 : c2);
    dp = dp > 90.0 ? 90.0 : (dp < -90.0 ? -90.0 : dp);
    double ap;
    const double zz = cosd(dp) * cosd(d0);
    if (fabs(zz) < 1e-10) {
        if (fabs(cosd(d0)) < 1e-10) ap = a0;
        else if (dp > 0) ap = a0 + phip - 180.0;
        else ap = a0 - phip;
    } else {
        const double xx = (0.0 - sind(dp) * sind(d0)) / zz;
        const double yy = sind(phip) / cosd(d0);
        ap = a0 - atan2(yy, xx) * kR2D;
    }
    w.alphap = ap;
    w.deltap = dp;
    w.phip = phip;
    w.sdp = sind(dp);
    w.cdp = cosd(dp);
    return w;
}

}  // namespace xts
