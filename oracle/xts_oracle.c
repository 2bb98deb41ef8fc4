/*
 * xts_oracle.c -- FP64 CPU restatement of the xtsrcmaps point-source
 * model-map path.  TEST INFRASTRUCTURE ONLY: this is the checker for the HIP
 * product path (see xts_oracle.h).  Citations are file:line in the reference
 * tree (Areustle/xtsrcmaps).
 *
 * Built with -ffp-contract=off so that every product and sum rounds exactly as
 * written.  Loops that the reference runs under `omp parallel for
 * schedule(static, 16)` over sources do the same here, so the timing of
 * or_cubature_convolve_map_psf is the reference algorithm's timing shape.
 */
#include "xts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Mathematical constants of the separation grid (misc/misc.hxx:11-39).  The
 * grid is sep[0] = 0, sep[i+1] = 1e-4 deg * exp(i * ln(7e5)/399). */
static const double PI        = 3.14159265358979323846264338327950288;
static const double TWOPI     = 6.28318530717958647692528676655900577;
static const double HALFPI    = 1.57079632679489661923132169163975144;
static const double D2R       = 0.01745329251994329576923690768488612713;
static const double R2D       = 57.2957795130823208767981548141051703324;
static const double SEP_STEP  = 0.033731417579011382769913057686378448039;
static const double RECIPSTEP = 29.645952402019046699874076387601854602589;
static const double REDM1     = 29.148763300179999028597222671574207003288;
static const double LN1E4     = 9.2103403719761827360719658187374568304044;

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Separations: psf/psf.hxx:24-31 (table), :33-43 (inverse), misc.hxx:22-39  */
/* ------------------------------------------------------------------------ */
void or_separations(double* sep) {
    const double xstep = log(70.0 / 1e-4) / 399.0;
    sep[0] = 0.0;
    for (int i = 0; i < 400; ++i) sep[i + 1] = 1e-4 * exp(i * xstep);
}

double or_sep_step(void) { return SEP_STEP; }

/* psf/psf.fast_separation_lower_index.cxx:5-22 -- note the threshold compares
 * the radius in degrees with 1 (not 1e-4), as coded. */
int or_fast_separation_lower_index(double v) {
    return v < 1. ? 1 : (int)floor(1. + log(1e4 * v) * RECIPSTEP);
}

/* ------------------------------------------------------------------------ */
/* lerp_pars: math/bilerp.cxx:6-27.  upper_bound bracket, weight of the     */
/* upper node, complement, and the upper index.  Returns -1 where the       */
/* reference would read outside the grid (an assert in Debug, UB otherwise). */
/* ------------------------------------------------------------------------ */
int or_lerp_pars(const double* g, long n, double v, int has_zlb, double zlb,
                 double* w, double* cw, long* idx) {
    if (has_zlb && v < zlb) {
        *w = 0.0; *cw = 0.0; *idx = 1;
        return 0;
    }
    long up = 0;
    while (up < n && !(v < g[up])) ++up; /* first element > v */
    if (up == 0 || up == n) return -1;
    double tt = (v - g[up - 1]) / (g[up] - g[up - 1]);
    *w = tt; *cw = 1. - tt; *idx = up;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* King PSF: irf/irf.hxx:22-45; analytic integral irf/irf.hxx:48-77         */
/* ------------------------------------------------------------------------ */
static double king_f(double u, double g) {
    if (g == 1.0) g = 1.001;
    return (1. - 1. / g) * pow(1. + u / g, -g);
}

double or_evaluate_king(double sep, double sf, const double* p) {
    const double ncore = p[0], ntail = p[1];
    const double score = p[2] * sf, stail = p[3] * sf;
    const double rc = sep / score, rt = sep / stail;
    const double uc = rc * rc / 2., ut = rt * rt / 2.;
    return ncore * king_f(uc, p[4]) + ntail * ncore * king_f(ut, p[5]);
}

static double base_f(double u, double g) {
    return 1. - pow(1. + u / g, 1. - g);
}

double or_psf_base_integral(double radius, double sf, const double* p) {
    const double ncore = p[0], ntail = p[1];
    const double score = p[2] * sf, stail = p[3] * sf;
    const double sep = radius * D2R;
    const double rc = sep / score, rt = sep / stail;
    const double uc = rc * rc / 2., ut = rt * rt / 2.;
    return ncore * base_f(uc, p[4]) * TWOPI * score * score
         + ntail * ncore * base_f(ut, p[5]) * TWOPI * stail * stail;
}

/* Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n with the
 * three-term recurrence (math/gauss_legendre.hxx:12-63 uses the same
 * textbook method, Abramowitz & Stegun 25.4.29). */
void or_legendre_rw(int n, double eps, double* x, double* w) {
    for (int i = 1; i <= n; ++i) {
        double z = cos(PI * (i - 0.25) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= n; ++j) {
                double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1; p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) <= eps) break;
        }
        x[i - 1] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* irf/irf.normalize_rpsf.cpp:11-61.  The scale products sp0, sp1 are formed
 * in float as the reference's IrfScale members are float. */
void or_normalize_rpsf(double* params, const double* logEs, long Mc, long Me,
                       float scale0, float scale1, float scale_index) {
    const float  sp0 = scale0 * scale0;
    const float  sp1 = scale1 * scale1;
    const double si  = scale_index;
    double gx[64], gw[64];
    or_legendre_rw(64, 1e-15, gx, gw);
    for (long c = 0; c < Mc; ++c) {
        for (long e = 0; e < Me; ++e) {
            double* p = params + (c * Me + e) * 6;
            const double energy = pow(10.0, logEs[e]);
            const double tt = pow(energy * 1.e-2, si);
            const double sf = sqrt((double)sp0 * tt * tt + (double)sp1);
            double norm;
            if (energy < 120.) {
                const double a = 0.0, b = 90.0;
                const double hw = 0.5 * (b - a), mid = 0.5 * (a + b);
                double acc = 0.0;
                for (int k = 0; k < 64; ++k) {
                    const double v = hw * gx[k] + mid;
                    const double kv = or_evaluate_king(v * D2R, sf, p);
                    const double y = sin(v * D2R) * TWOPI * D2R;
                    acc += gw[k] * (kv * y);
                }
                norm = acc * hw;
            } else {
                norm = or_psf_base_integral(90.0, sf, p);
            }
            p[0] /= norm;
            p[2] *= sf;
            p[3] *= sf;
            p[4] = p[4] == 1. ? 1.001 : p[4];
            p[5] = p[5] == 1. ? 1.001 : p[5];
        }
    }
}

/* irf/irf.prepare_grid.cpp:26-102: bin centres, padding with -1/+1 in cos
 * theta and 0/10 in log10 E, and edge duplication of the parameter block. */
void or_prepare_grid(const float* e_lo, const float* e_hi, long Me0,
                     const float* c_lo, const float* c_hi, long Mc0,
                     const float* raw, long Ng, double* cosths, double* logEs,
                     double* params, double* min_costheta) {
    const long Mc = Mc0 + 2, Me = Me0 + 2;
    for (long i = 0; i < Mc0; ++i) {
        const float s = c_lo[i] + c_hi[i];
        cosths[i + 1] = 0.5 * (double)s;
    }
    cosths[0] = -1.;
    cosths[Mc - 1] = 1.;
    for (long i = 0; i < Me0; ++i) {
        const float pr = e_lo[i] * e_hi[i];
        const float lg = log10f(pr);
        logEs[i + 1] = 0.5 * (double)lg;
    }
    logEs[0] = 0.;
    logEs[Me - 1] = 10.;
    for (long p = 0; p < Ng; ++p)
        for (long t = 0; t < Mc; ++t) {
            const long t_ = t == 0 ? 0 : (t >= Mc0 ? Mc0 - 1 : t - 1);
            for (long e = 0; e < Me; ++e) {
                const long e_ = e == 0 ? 0 : (e >= Me0 ? Me0 - 1 : e - 1);
                params[(t * Me + e) * Ng + p] = raw[(p * Mc0 + t_) * Me0 + e_];
            }
        }
    *min_costheta = c_lo[0];
}

/* irf/irf.cxx:70-109.  Row 1 comes from p0 and row 0 from p1, as coded; the
 * break intercepts b1, b2 are formed in float as the inputs are float. */
static void ltf_row(const float* p, const double* logEs, long E, double* out) {
    const float a0 = p[0], b0 = p[1], a1 = p[2], lb1 = p[3], a2 = p[4], lb2 = p[5];
    const float b1f = (a0 - a1) * lb1 + b0;
    const float b2f = (a1 - a2) * lb2 + b1f;
    const double b1 = b1f, b2 = b2f;
    for (long e = 0; e < E; ++e) {
        const double x = logEs[e];
        out[e] = x < lb1 ? a0 * x + b0 : (x < lb2 ? a1 * x + b1 : a2 * x + b2);
    }
}

void or_ltf(const double* logEs, long E, const float* p0, const float* p1,
            double* out) {
    ltf_row(p0, logEs, E, out + E);
    ltf_row(p1, logEs, E, out);
}

/* irf/irf.aeff_value.cxx:11-55: bilinear in (cos theta, log E), zero below
 * minCosTheta (lerp weights 0), times 1e4 (m^2 -> cm^2). */
int or_aeff_value(const double* costh, long C, const double* logE, long E,
                  const double* pc, long Mc, const double* pe, long Me,
                  const double* params, double minc, double* out) {
    for (long c = 0; c < C; ++c) {
        double cw, ccw; long ci;
        if (or_lerp_pars(pc, Mc, costh[c], 1, minc, &cw, &ccw, &ci)) return -1;
        for (long e = 0; e < E; ++e) {
            double ew, ecw; long ei;
            if (or_lerp_pars(pe, Me, logE[e], 0, 0., &ew, &ecw, &ei)) return -1;
            const double xx = ccw * ecw * params[(ci - 1) * Me + ei - 1];
            const double xy = ccw * ew * params[(ci - 1) * Me + ei];
            const double yx = cw * ecw * params[ci * Me + ei - 1];
            const double yy = cw * ew * params[ci * Me + ei];
            out[c * E + e] = 1e4 * (xx + xy + yx + yy);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* (a) Exposure                                                             */
/* ------------------------------------------------------------------------ */
/* exposure/exposure.cxx:29-40 */
void or_costhetas(long nbins, double cosmin, int thetabin, double* out) {
    for (long i = 0; i < nbins; ++i) {
        double f = (i + 0.5) / nbins;
        f = thetabin ? f * f : f;
        out[i] = 1. - f * (1. - cosmin);
    }
}

/* HEALPix NESTED ang2pix (Gorski et al. 2005, ApJ 622:759; the reference
 * follows the same published algorithm in healpix/healpix.cxx:182-229). */
static double fmodulo(double v, double m) {
    if (v >= 0) return v < m ? v : fmod(v, m);
    double r = fmod(v, m) + m;
    return r == m ? 0.0 : r;
}

static uint64_t interleave_xy(uint64_t ix, uint64_t iy) {
    uint64_t r = 0;
    for (int b = 0; b < 32; ++b) {
        r |= ((ix >> b) & 1ULL) << (2 * b);
        r |= ((iy >> b) & 1ULL) << (2 * b + 1);
    }
    return r;
}

uint64_t or_ang2pix_nest(double theta, double phi, int64_t nside) {
    int order = 0;
    while ((1LL << order) < nside) ++order;
    const int have_sth = (theta < 0.01) || (theta > PI - 0.01);
    const double sth = have_sth ? sin(theta) : 0.;
    const double z = cos(theta), za = fabs(z);
    const double tt = fmodulo(phi * (2.0 / PI), 4.0);
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
        const double tmp = ((za < 0.99) || !have_sth)
                               ? nside * sqrt(3 * (1 - za))
                               : nside * sth / sqrt((1. + za) / 3.);
        int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
        if (jp > nside - 1) jp = nside - 1;
        if (jm > nside - 1) jm = nside - 1;
        if (z >= 0) { face = ntt; ix = nside - jm - 1; iy = nside - jp - 1; }
        else        { face = ntt + 8; ix = jp; iy = jm; }
    }
    return (face << (2 * order)) + interleave_xy(ix, iy);
}

/* exposure/exposure.cxx:42-63: copy the livetime row of the NESTED pixel
 * containing each source (theta = pi/2 - dec, phi = ra). */
void or_src_cosbins(const double* src_sph, long S, const double* cube,
                    int64_t nside, long C, double* out) {
    for (long s = 0; s < S; ++s) {
        const double theta = HALFPI - src_sph[2 * s + 1] * D2R;
        const double phi = D2R * src_sph[2 * s];
        const uint64_t pix = or_ang2pix_nest(theta, phi, nside);
        memcpy(out + s * C, cube + pix * C, (size_t)C * sizeof(double));
    }
}

/* exposure/exposure.contract.cpp:13-42: C[m,n] = sum_k A[m,k] B[k,n]. */
void or_contract(const double* A, const double* B, long M, long K, long N,
                 double* out) {
    for (long m = 0; m < M; ++m)
        for (long n = 0; n < N; ++n) {
            double acc = 0.0;
            for (long k = 0; k < K; ++k) acc += A[m * K + k] * B[k * N + n];
            out[m * N + n] = acc;
        }
}

/* exposure/exposure.cxx:69-173: four contractions, LTF row scaling (the
 * back section also uses the front LTF), and the sum
 * ((lef + leb) + lwf) + lwb in that order. */
void or_exposure(const double* ecb, const double* wecb, const double* af,
                 const double* ab, const double* ltf, long S, long C, long E,
                 double* out, double* lef, double* lwf, double* leb,
                 double* lwb) {
    const size_t n = (size_t)S * E;
    double* t = (double*)malloc(4 * n * sizeof(double));
    double *ef = t, *wf = t + n, *eb = t + 2 * n, *wb = t + 3 * n;
    or_contract(ecb, af, S, C, E, ef);
    or_contract(wecb, af, S, C, E, wf);
    or_contract(ecb, ab, S, C, E, eb);
    or_contract(wecb, ab, S, C, E, wb);
    for (long s = 0; s < S; ++s)
        for (long e = 0; e < E; ++e) {
            const size_t i = (size_t)s * E + e;
            ef[i] *= ltf[e];
            wf[i] *= ltf[E + e];
            eb[i] *= ltf[e];
            wb[i] *= ltf[E + e];
        }
    if (lef) memcpy(lef, ef, n * sizeof(double));
    if (lwf) memcpy(lwf, wf, n * sizeof(double));
    if (leb) memcpy(leb, eb, n * sizeof(double));
    if (lwb) memcpy(lwb, wb, n * sizeof(double));
    for (size_t i = 0; i < n; ++i) out[i] = ((ef[i] + eb[i]) + wf[i]) + wb[i];
    free(t);
}

/* ------------------------------------------------------------------------ */
/* (b) Mean PSF                                                             */
/* ------------------------------------------------------------------------ */
/* psf/psf.king.cxx:9-43: K[d, c, e] at the 401 separations (radians). */
void or_king(const double* params, long Mc, long Me, double* out) {
    double sep[OR_ND];
    or_separations(sep);
    for (long d = 0; d < OR_ND; ++d)
        for (long c = 0; c < Mc; ++c)
            for (long e = 0; e < Me; ++e)
                out[(d * Mc + c) * Me + e] =
                    or_evaluate_king(sep[d] * D2R, 1.0, params + (c * Me + e) * 6);
}

/* psf/psf.bilerp.cxx:6-52: [C, D, E] bilinear samples of the King table. */
int or_bilerp(const double* costh, long C, const double* logE, long E,
              const double* pc, long Mc, const double* pe, long Me,
              const double* K, double* out) {
    for (long c = 0; c < C; ++c) {
        double cw, ccw; long ci;
        if (or_lerp_pars(pc, Mc, costh[c], 0, 0., &cw, &ccw, &ci)) return -1;
        for (long e = 0; e < E; ++e) {
            double ew, ecw; long ei;
            if (or_lerp_pars(pe, Me, logE[e], 0, 0., &ew, &ecw, &ei)) return -1;
            for (long d = 0; d < OR_ND; ++d) {
                const double* Kd = K + d * Mc * Me;
                out[(c * OR_ND + d) * E + e] =
                    ccw * ecw * Kd[(ci - 1) * Me + ei - 1] +
                    ccw * ew * Kd[(ci - 1) * Me + ei] +
                    cw * ecw * Kd[ci * Me + ei - 1] +
                    cw * ew * Kd[ci * Me + ei];
            }
        }
    }
    return 0;
}

/* psf/psf.corrected_exposure_psf.cxx:12-120:
 * (ECB . (psf o aeff)) * LTF[0] + (WECB . (psf o aeff)) * LTF[1]. */
void or_corrected_exposure_psf(const double* obs_psf, const double* aeff,
                               const double* ecb, const double* wecb,
                               const double* ltf, long S, long C, long D,
                               long E, double* out) {
    const long N = D * E;
    double* pa = (double*)malloc((size_t)C * N * sizeof(double));
    for (long c = 0; c < C; ++c)
        for (long d = 0; d < D; ++d)
            for (long e = 0; e < E; ++e)
                pa[c * N + d * E + e] = obs_psf[c * N + d * E + e] * aeff[c * E + e];
    double* w = (double*)malloc((size_t)S * N * sizeof(double));
    or_contract(ecb, pa, S, C, N, out);
    or_contract(wecb, pa, S, C, N, w);
    for (long s = 0; s < S; ++s)
        for (long d = 0; d < D; ++d)
            for (long e = 0; e < E; ++e) {
                const size_t i = (size_t)s * N + d * E + e;
                out[i] *= ltf[e];
                w[i] *= ltf[E + e];
                out[i] = out[i] + w[i];
            }
    free(w);
    free(pa);
}

/* psf/psf.mean_psf.cxx:6-33: (front + back) / exposure[s, e]. */
void or_mean_psf(const double* front, const double* back, const double* expo,
                 long S, long D, long E, double* out) {
    for (long s = 0; s < S; ++s)
        for (long d = 0; d < D; ++d)
            for (long e = 0; e < E; ++e) {
                const size_t i = ((size_t)s * D + d) * E + e;
                out[i] = (front[i] + back[i]) / expo[s * E + e];
            }
}

/* psf/psf.partial_total_integral.cxx:8-52: per (s, e), exact integral of
 * the piecewise-linear 2 pi sin(theta) psf(theta) over each separation
 * segment, cumulated, then normalised by the total when it is non-zero. */
void or_partial_total_integral(const double* upsf, long S, long D, long E,
                               double* parint, double* totint) {
    double sep[OR_ND];
    or_separations(sep);
    for (long s = 0; s < S; ++s)
        for (long e = 0; e < E; ++e) {
            const double* u = upsf + (size_t)s * D * E + e;
            double* pi = parint + (size_t)s * D * E + e;
            double cum = 0.0;
            pi[0] = 0.0;
            for (long d = 1; d < D; ++d) {
                const double t1 = D2R * sep[d - 1], t2 = D2R * sep[d];
                const double y1 = TWOPI * u[(d - 1) * E] * sin(t1);
                const double y2 = TWOPI * u[d * E] * sin(t2);
                const double m = (y2 - y1) / (t2 - t1);
                const double b = y1 - t1 * m;
                const double v = 0.5 * m * (t2 * t2 - t1 * t1) + b * (t2 - t1);
                cum += v;
                pi[d * E] = cum;
            }
            totint[s * E + e] = cum;
            if (cum != 0.0)
                for (long d = 0; d < D; ++d) pi[d * E] /= cum;
        }
}

/* psf/psf.normalize.cxx:7-24 */
void or_normalize(double* upsf, const double* tot, long S, long D, long E) {
    for (long s = 0; s < S; ++s)
        for (long e = 0; e < E; ++e)
            for (long d = 0; d < D; ++d)
                upsf[((size_t)s * D + d) * E + e] /= tot[s * E + e];
}

/* psf/psf.psf_lookup_table_and_partial_integrals.cxx:8-57 (back uses the
 * front LTF, :42-47). */
int or_psf_lookup_table_and_partial_integrals(
    const double* params_f, const double* pc_f, long Mc_f, const double* pe_f,
    long Me_f, const double* params_b, const double* pc_b, long Mc_b,
    const double* pe_b, long Me_b, const double* costh, long C,
    const double* logE, long E, const double* af, const double* ab,
    const double* ecb, const double* wecb, const double* ltf,
    const double* expo, long S, double* upsf, double* parint) {
    const long D = OR_ND;
    double* kf = (double*)malloc((size_t)D * Mc_f * Me_f * sizeof(double));
    double* kb = (double*)malloc((size_t)D * Mc_b * Me_b * sizeof(double));
    double* bf = (double*)malloc((size_t)C * D * E * sizeof(double));
    double* bb = (double*)malloc((size_t)C * D * E * sizeof(double));
    double* cf = (double*)malloc((size_t)S * D * E * sizeof(double));
    double* cb = (double*)malloc((size_t)S * D * E * sizeof(double));
    double* tot = (double*)malloc((size_t)S * E * sizeof(double));
    int rc = 0;
    or_king(params_f, Mc_f, Me_f, kf);
    or_king(params_b, Mc_b, Me_b, kb);
    rc |= or_bilerp(costh, C, logE, E, pc_f, Mc_f, pe_f, Me_f, kf, bf);
    rc |= or_bilerp(costh, C, logE, E, pc_b, Mc_b, pe_b, Me_b, kb, bb);
    if (rc == 0) {
        or_corrected_exposure_psf(bf, af, ecb, wecb, ltf, S, C, D, E, cf);
        or_corrected_exposure_psf(bb, ab, ecb, wecb, ltf, S, C, D, E, cb);
        or_mean_psf(cf, cb, expo, S, D, E, upsf);
        or_partial_total_integral(upsf, S, D, E, parint, tot);
        or_normalize(upsf, tot, S, D, E);
    }
    free(kf); free(kb); free(bf); free(bb); free(cf); free(cb); free(tot);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* WCS.  sky_geom/sky_geom.hxx:79-125 sets CTYPE, CRPIX, CRVAL (crval1>180 */
/* -> -360), CDELT and CROTA2 (altlin |= 4) in a wcslib wcsprm.  Restated  */
/* from the FITS WCS papers (Greisen & Calabretta 2002, Calabretta &       */
/* Greisen 2002) for the CAR and AIT projections, default LONPOLE/LATPOLE. */
/* ------------------------------------------------------------------------ */
static double sind(double a) { return sin(a * D2R); }
static double cosd(double a) { return cos(a * D2R); }
static double atan2d(double y, double x) { return atan2(y, x) * R2D; }
static double asind(double v) { return asin(v) * R2D; }
static double acosd(double v) { return acos(v) * R2D; }

void or_wcs_setup(or_wcs* w) {
    /* CROTA2 -> PC matrix (Paper I eq 189-190 convention as wcslib uses):
     * x = cdelt1 cos(r) dp1 - cdelt2 sin(r) dp2
     * y = cdelt1 sin(r) dp1 + cdelt2 cos(r) dp2 */
    const double r = w->crota2;
    const double cr = cosd(r), sr = sind(r);
    w->m11 = w->cdelt1 * cr;  w->m12 = -w->cdelt2 * sr;
    w->m21 = w->cdelt1 * sr;  w->m22 = w->cdelt2 * cr;
    const double det = w->m11 * w->m22 - w->m12 * w->m21;
    w->i11 = w->m22 / det;  w->i12 = -w->m12 / det;
    w->i21 = -w->m21 / det; w->i22 = w->m11 / det;

    /* Celestial frame: native reference point (phi0, theta0) = (0, 0) for
     * both CAR and AIT.  Default LONPOLE = 0 if delta0 >= theta0 else 180,
     * LATPOLE = +90 (Paper II sect. 2.4). */
    const double a0 = w->crval1 > 180 ? w->crval1 - 360 : w->crval1;
    const double d0 = w->crval2;
    const double phi0 = 0.0, theta0 = 0.0, latpole = 90.0;
    const double phip = d0 >= theta0 ? 0.0 : 180.0;
    double dp;
    {
        const double x = cosd(theta0) * cosd(phip - phi0), y = sind(theta0);
        const double z = sqrt(x * x + y * y);
        const double u = atan2d(y, x);
        const double v = acosd(sind(d0) / z);
        double c1 = u + v, c2 = u - v;
        if (c1 > 180) c1 -= 360; else if (c1 < -180) c1 += 360;
        if (c2 > 180) c2 -= 360; else if (c2 < -180) c2 += 360;
        const int ok1 = fabs(c1) <= 90.0 + 1e-10, ok2 = fabs(c2) <= 90.0 + 1e-10;
        if (ok1 && ok2) dp = fabs(c1 - latpole) < fabs(c2 - latpole) ? c1 : c2;
        else dp = ok1 ? c1 : c2;
        if (dp > 90) dp = 90; else if (dp < -90) dp = -90;
    }
    double ap;
    const double zz = cosd(dp) * cosd(d0);
    if (fabs(zz) < 1e-10) {
        if (fabs(cosd(d0)) < 1e-10) ap = a0;
        else if (dp > 0) ap = a0 + phip - phi0 - 180.0;
        else ap = a0 - phip + phi0;
    } else {
        const double x = (sind(theta0) - sind(dp) * sind(d0)) / zz;
        const double y = sind(phip - phi0) * cosd(theta0) / cosd(d0);
        ap = a0 - atan2d(y, x);
    }
    w->alphap = ap;
    w->deltap = dp;
    w->phip = phip;
}

/* native (phi, theta) -> celestial (lon, lat), Paper II eq 2 */
static void nat2cel(const or_wcs* w, double phi, double theta, double* lon,
                    double* lat) {
    const double dphi = phi - w->phip;
    const double st = sind(theta), ct = cosd(theta);
    const double sdp = sind(w->deltap), cdp = cosd(w->deltap);
    const double x = st * cdp - ct * sdp * cosd(dphi);
    const double y = -ct * sind(dphi);
    double a = w->alphap + atan2d(y, x);
    double s = st * sdp + ct * cdp * cosd(dphi);
    if (s > 1) s = 1; else if (s < -1) s = -1;
    *lon = a;
    *lat = asind(s);
}

/* celestial -> native, Paper II eq 5 */
static void cel2nat(const or_wcs* w, double lon, double lat, double* phi,
                    double* theta) {
    const double da = lon - w->alphap;
    const double sd = sind(lat), cd = cosd(lat);
    const double sdp = sind(w->deltap), cdp = cosd(w->deltap);
    const double x = sd * cdp - cd * sdp * cosd(da);
    const double y = -cd * sind(da);
    double p = w->phip + atan2d(y, x);
    while (p > 180) p -= 360;
    while (p < -180) p += 360;
    double s = sd * sdp + cd * cdp * cosd(da);
    if (s > 1) s = 1; else if (s < -1) s = -1;
    *phi = p;
    *theta = asind(s);
}

int or_pix2sph(const or_wcs* w, double p1, double p2, double* lon, double* lat) {
    const double dp1 = p1 - w->crpix1, dp2 = p2 - w->crpix2;
    const double x = w->m11 * dp1 + w->m12 * dp2;
    const double y = w->m21 * dp1 + w->m22 * dp2;
    double phi, theta;
    if (w->proj == OR_PROJ_CAR) {
        phi = x; theta = y;
        if (fabs(theta) > 90.0) return -1;
    } else {
        /* AIT deprojection, Paper II eq 105-108 with R0 = 180/pi */
        const double r0 = R2D;
        const double u = 1.0 - x * x / (16.0 * r0 * r0) - y * y / (4.0 * r0 * r0);
        if (u < 0.0) return -1;
        const double z = sqrt(u);
        double s = z * y / r0;
        if (fabs(s) > 1.0) return -1;
        theta = asind(s);
        const double xp = 2.0 * z * z - 1.0, yp = z * x / (2.0 * r0);
        phi = (xp == 0.0 && yp == 0.0) ? 0.0 : 2.0 * atan2d(yp, xp);
    }
    double a, d;
    nat2cel(w, phi, theta, &a, &d);
    while (a < 0) a += 360.;
    while (a >= 360) a -= 360.;
    *lon = a;
    *lat = d;
    return 0;
}

int or_sph2pix(const or_wcs* w, double lon, double lat, double* p1, double* p2) {
    if (lon > 180) lon -= 360.;
    if (lon < -180) lon += 360.;
    double phi, theta, x, y;
    cel2nat(w, lon, lat, &phi, &theta);
    if (w->proj == OR_PROJ_CAR) {
        x = phi; y = theta;
    } else {
        const double r0 = R2D;
        const double ct = cosd(theta), st = sind(theta);
        const double g = sqrt(2.0 * r0 * r0 / (1.0 + ct * cosd(0.5 * phi)));
        x = 2.0 * g * ct * sind(0.5 * phi);
        y = g * st;
    }
    *p1 = w->crpix1 + w->i11 * x + w->i12 * y;
    *p2 = w->crpix2 + w->i21 * x + w->i22 * y;
    return 0;
}

/* sky_geom/sky_geom.hxx:281-287 */
void or_sph2dir(double lon, double lat, double* d) {
    const double ca = cos(lon * D2R), cd = cos(lat * D2R);
    const double sa = sin(lon * D2R), sd = sin(lat * D2R);
    d[0] = ca * cd; d[1] = sa * cd; d[2] = sd;
}

int or_pix2dir(const or_wcs* w, double p1, double p2, double* d) {
    double lon, lat;
    int rc = or_pix2sph(w, p1, p2, &lon, &lat);
    or_sph2dir(lon, lat, d);
    return rc;
}

/* sky_geom/sky_geom.hxx:305-310 */
double or_dir_diff(const double* a, const double* b) {
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return 2. * asin(0.5 * sqrt(x * x + y * y + z * z));
}

/* ------------------------------------------------------------------------ */
/* (c) Model map                                                            */
/* ------------------------------------------------------------------------ */
/* mm_cubature/points_weights.hpp:5-56: per pixel, the four corner unit
 * vectors at 1-based pixel coordinates (h+1 +- 1/2, w+1 +- 1/2), a bilinear
 * (un-normalised) blend at each rule point (u along w, v along h) and the
 * weight w/4.  out is [Nh, Nw, 4N] of (x, y, z, weight). */
void or_square_ptswts(long Nh, long Nw, const or_wcs* wc, const double* rule,
                      long N, double* out) {
    for (long h = 0; h < Nh; ++h)
        for (long w = 0; w < Nw; ++w) {
            const double ph = 1. + h, pw = 1. + w;
            double a[3], b[3], c[3], d[3];
            or_pix2dir(wc, ph - 0.5, pw - 0.5, a);
            or_pix2dir(wc, ph - 0.5, pw + 0.5, b);
            or_pix2dir(wc, ph + 0.5, pw + 0.5, c);
            or_pix2dir(wc, ph + 0.5, pw - 0.5, d);
            double* o = out + (h * Nw + w) * 4 * N;
            for (long p = 0; p < N; ++p) {
                const double u = rule[3 * p], v = rule[3 * p + 1];
                const double nn = (1.0 - u) * (1.0 - v);
                const double pn = (1.0 + u) * (1.0 - v);
                const double pp = (1.0 + u) * (1.0 + v);
                const double np = (1.0 - u) * (1.0 + v);
                for (int k = 0; k < 3; ++k)
                    o[4 * p + k] =
                        0.25 * (nn * a[k] + pn * b[k] + pp * c[k] + np * d[k]);
                o[4 * p + 3] = 0.25 * rule[3 * p + 2];
            }
        }
}

/* mm_cubature/weights.naive.4096.cpp:8-61 generalised to npix x npix
 * midpoints with exact (projected) directions and weight 1/npix^2. */
void or_naive_ptswts(long Nh, long Nw, const or_wcs* wc, long npix, double* out) {
    const double dstep = 1.0 / (double)npix;
    const double wt = 1.0 / (double)(npix * npix);
    for (long h = 0; h < Nh; ++h)
        for (long w = 0; w < Nw; ++w) {
            double* o = out + (h * Nw + w) * 4 * npix * npix;
            for (long p = 0; p < npix * npix; ++p) {
                const long i = p / npix, j = p % npix;
                const double x = 1. + h + (-0.5 + (i + 0.5) * dstep);
                const double y = 1. + w + (-0.5 + (j + 0.5) * dstep);
                or_pix2dir(wc, x, y, o + 4 * p);
                o[4 * p + 3] = wt;
            }
        }
}

/* Grouped accumulation shared by both pixel kernels: per distinct upper
 * index the interpolation weights are summed first, then each group adds
 * va L[i-1, e] + vb L[i, e] -- the reference's collapsed-point scheme
 * (convolve_pixel_psf_logsep.hpp:56-95, convolve_pixel_psf_searchsep.hpp:
 * 41-73). */
typedef struct { long idx; double va, vb; } or_group;

static void flush_groups(double* map, long E, const double* lut,
                         const or_group* g, long ng) {
    for (long k = 0; k < ng; ++k) {
        const double* L = lut + (size_t)(g[k].idx - 1) * E;
        const double* R = lut + (size_t)g[k].idx * E;
        for (long e = 0; e < E; ++e) map[e] += (g[k].va * L[e]) + (g[k].vb * R[e]);
    }
}

static long add_group(or_group* g, long ng, long idx, double va, double vb,
                      int sorted) {
    long k = 0;
    for (; k < ng; ++k) if (g[k].idx == idx) break;
    if (k < ng) { g[k].va += va; g[k].vb += vb; return ng; }
    if (!sorted) { g[ng].idx = idx; g[ng].va = va; g[ng].vb = vb; return ng + 1; }
    long at = 0;
    while (at < ng && g[at].idx < idx) ++at;
    for (long m = ng; m > at; --m) g[m] = g[m - 1];
    g[at].idx = idx; g[at].va = va; g[at].vb = vb;
    return ng + 1;
}

static double point_sep_deg(const double* pw, const double* src) {
    const double x = pw[0] - src[0], y = pw[1] - src[1], z = pw[2] - src[2];
    return 2.0 * R2D * asin(0.5 * sqrt(x * x + y * y + z * z));
}

/* mm_cubature/convolve_pixel_psf_logsep.hpp:23-96 with its intended
 * semantics (the grouping keys are stored; as shipped they never are):
 * map[e] += sum over groups of va L[i-1, e] + vb L[i, e], with the
 * reference's index and fraction formulas, including the i = 2 linear
 * extrapolation below 1e-4 deg and the clamp to i = 400 above 70 deg. */
void or_pixel_logsep(double* map, long E, const double* pw, long N,
                     const double* lut, const double* src) {
    or_group* g = (or_group*)malloc((size_t)N * sizeof(or_group));
    long ng = 0;
    for (long p = 0; p < N; ++p) {
        const double w = pw[4 * p + 3];
        const double s = point_sep_deg(pw + 4 * p, src);
        const double j = s < 1e-4 ? 1. : 1. + RECIPSTEP * (LN1E4 + log(s));
        unsigned idx = 1u + (unsigned)(uint16_t)j;
        if (idx > 400u) idx = 400u;
        const double t = REDM1 * (s * 1e4 * exp(SEP_STEP * (2. - (double)idx)) - 1.);
        ng = add_group(g, ng, (long)idx, (1.0 - t) * w, t * w, 0);
    }
    flush_groups(map, E, lut, g, ng);
    free(g);
}

/* mm_cubature/convolve_pixel_psf_searchsep.hpp:26-75: the bracketing node
 * found by upper_bound on the true separations, groups in ascending index
 * order (std::map).  Used for the 4096-point midpoint fixture. */
void or_pixel_searchsep(double* map, long E, const double* pw, long N,
                        const double* lut, const double* src) {
    double sep[OR_ND];
    or_separations(sep);
    or_group* g = (or_group*)malloc((size_t)N * sizeof(or_group));
    long ng = 0;
    for (long p = 0; p < N; ++p) {
        const double w = pw[4 * p + 3];
        const double s = point_sep_deg(pw + 4 * p, src);
        long lo = 0, hi = OR_ND;
        while (lo < hi) { long m = (lo + hi) / 2; if (sep[m] <= s) lo = m + 1; else hi = m; }
        long idx = lo;
        if (idx >= OR_ND) idx = OR_ND - 1; /* the reference reads past the end here */
        const double t = (s - sep[idx - 1]) / (sep[idx] - sep[idx - 1]);
        ng = add_group(g, ng, idx, (1.0 - t) * w, t * w, 1);
    }
    flush_groups(map, E, lut, g, ng);
    free(g);
}

/* mm_cubature/cubature.implementation.cxx:21-55: out [S, H, W, E], zeroed,
 * loop s (OpenMP static,16) -> w -> h.  ptswts is [Nh, Nw, 4N]. */
void or_cubature_convolve_map_psf(long Nh, long Nw, const double* src_sph,
                                  long S, const double* lut, long E,
                                  const or_wcs* wc, const double* ptswts,
                                  long N, int mode, double* out) {
    (void)wc;
    memset(out, 0, (size_t)S * Nh * Nw * E * sizeof(double));
    long s;
#pragma omp parallel for schedule(static, 16)
    for (s = 0; s < S; ++s) {
        double src[3];
        or_sph2dir(src_sph[2 * s], src_sph[2 * s + 1], src);
        const double* L = lut + (size_t)s * OR_ND * E;
        for (long w = 0; w < Nw; ++w)
            for (long h = 0; h < Nh; ++h) {
                double* m = out + (((size_t)s * Nh + h) * Nw + w) * E;
                const double* pw = ptswts + (h * Nw + w) * 4 * N;
                if (mode == OR_RULE_SEARCHSEP) or_pixel_searchsep(m, E, pw, N, L, src);
                else or_pixel_logsep(m, E, pw, N, L, src);
            }
    }
}

/* math/solid_angle.hpp:42-79: the quad of pixel centres (h+1, w+1) ..
 * (h+2, w+2) in 1-based pixel coordinates, split in two triangles. */
static double vnorm(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double tri_omega(const double* A, const double* B, const double* D) {
    double u[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double v[3] = {A[0] - D[0], A[1] - D[1], A[2] - D[2]};
    const double nu = vnorm(u), nv = vnorm(v);
    for (int k = 0; k < 3; ++k) { u[k] /= nu; v[k] /= nv; }
    double c[3];
    cross3(u, v, c);
    return or_dir_diff(A, B) * or_dir_diff(A, D) * vnorm(c);
}

void or_solid_angle(long Nh, long Nw, const or_wcs* wc, double* out) {
    for (long h = 0; h < Nh; ++h) {
        const double ph = 1. + h;
        for (long w = 0; w < Nw; ++w) {
            const double pw = 1. + w;
            double A[3], B[3], C[3], D[3];
            or_pix2dir(wc, ph, pw, A);
            or_pix2dir(wc, ph, pw + 1., B);
            or_pix2dir(wc, ph + 1., pw + 1., C);
            or_pix2dir(wc, ph + 1., pw, D);
            out[h * Nw + w] = 0.5 * (tri_omega(A, B, D) + tri_omega(C, B, D));
        }
    }
}

/* model_map/model_map.psf_boundary_radius.cxx:9-62: in-FOV test with a
 * 3.5 pixel buffer, and the minimum distance (deg) from the source to the
 * 4 (N+1) half-integer boundary points. */
void or_psf_boundary_radius(long Nh, long Nw, const double* src_sph, long S,
                            const or_wcs* wc, double* radius, uint8_t* in_fov) {
    const double buf = 3.5;
    for (long s = 0; s < S; ++s) {
        double src[3], p1, p2, d[3];
        or_sph2dir(src_sph[2 * s], src_sph[2 * s + 1], src);
        or_sph2pix(wc, src_sph[2 * s], src_sph[2 * s + 1], &p1, &p2);
        in_fov[s] = p1 > buf && p1 < (Nh - buf) && p2 > buf && p2 < (Nw - buf);
        double mn = 360.;
        for (long h = 0; h <= Nh; ++h) {
            or_pix2dir(wc, (float)h + 0.5f, 0.5f, d);
            double v = R2D * or_dir_diff(src, d);
            mn = mn < v ? mn : v;
            or_pix2dir(wc, (float)h + 0.5f, (float)Nw + 0.5f, d);
            v = R2D * or_dir_diff(src, d);
            mn = mn < v ? mn : v;
        }
        for (long w = 0; w <= Nw; ++w) {
            or_pix2dir(wc, 0.5f, (float)w + 0.5f, d);
            double v = R2D * or_dir_diff(src, d);
            mn = mn < v ? mn : v;
            or_pix2dir(wc, (float)Nh + 0.5f, (float)w + 0.5f, d);
            v = R2D * or_dir_diff(src, d);
            mn = mn < v ? mn : v;
        }
        radius[s] = mn;
    }
}

/* model_map/model_map.map_integral.cxx:4-48: sum of the (solid-angle
 * scaled) map over pixels whose centre lies within the radius, then the
 * reciprocal (0 stays 0).  map is [S, H, W, E]. */
void or_map_integral(const double* map, long S, long Nh, long Nw, long E,
                     const double* src_sph, const or_wcs* wc,
                     const double* radius, const uint8_t* in_fov, double* out) {
    memset(out, 0, (size_t)S * E * sizeof(double));
    long s;
#pragma omp parallel for schedule(static, 16)
    for (s = 0; s < S; ++s) {
        if (!in_fov[s]) continue;
        double src[3], d[3];
        or_sph2dir(src_sph[2 * s], src_sph[2 * s + 1], src);
        for (long h = 0; h < Nh; ++h)
            for (long w = 0; w < Nw; ++w) {
                or_pix2dir(wc, (float)h + 1.0f, (float)w + 1.0f, d);
                const double rad = or_dir_diff(src, d);
                if (rad <= D2R * radius[s]) {
                    const double* m = map + (((size_t)s * Nh + h) * Nw + w) * E;
                    for (long e = 0; e < E; ++e) out[s * E + e] += m[e];
                }
            }
    }
    for (long i = 0; i < S * E; ++i) out[i] = out[i] != 0.0 ? 1.0 / out[i] : out[i];
}

/* The PSF integral to the correction radius of each (source, energy).
 * mode 0: model_map/model_map.map_correction_factor.cxx:31-47 as coded: the
 *   node index d of psf/psf.fast_separation_lower_index.cxx:5-22 (1 for
 *   radii below 1 degree, Q5), clamped to 399, and the whole segment
 *   d -> d+1 of the linear interpolant of 2 pi sin(theta) psf (Q6).
 * mode 1: the FermiTools-consistent variant: d the node below the radius
 *   (sep[d] <= R < sep[d+1]) and the segment integrated to exactly R.
 * Not capped (the caller caps at 1). */
void or_psf_radius_integral(const double* radius, const double* upsf,
                            const double* parint, long S, long D, long E,
                            int mode, double* out) {
    double sep[OR_ND];
    or_separations(sep);
    for (long s = 0; s < S; ++s) {
        const double R = radius[s];
        long d;
        if (mode == 0) {
            d = or_fast_separation_lower_index(R);
        } else {
            d = R < 1e-4 ? 0 : (long)floor(1.0 + log(1e4 * R) * RECIPSTEP);
            while (d + 1 < OR_ND && sep[d + 1] <= R) ++d;   /* rounding at a node */
            while (d > 0 && sep[d] > R) --d;
        }
        if (d >= OR_ND - 2) d = OR_ND - 2;
        const double t1 = sep[d] * D2R, t2 = sep[d + 1] * D2R;
        const double te = mode == 0 ? t2 : R * D2R;
        for (long e = 0; e < E; ++e) {
            const double y1 = TWOPI * upsf[((size_t)s * D + d) * E + e] * sin(t1);
            const double y2 = TWOPI * upsf[((size_t)s * D + d + 1) * E + e] * sin(t2);
            const double m = (y2 - y1) / (t2 - t1);
            const double b = y1 - t1 * m;
            const double v = 0.5 * m * (te * te - t1 * t1) + b * (te - t1);
            out[s * E + e] = v + parint[((size_t)s * D + d) * E + e];
        }
    }
}

/* model_map/model_map.map_correction_factor.cxx:3-53 (mode 0, as coded) or
 * its FermiTools-consistent variant (mode 1): the PSF integral to the
 * radius, capped at 1, times the inverse map integral; 1 outside FOV. */
void or_map_correction_factor_mode(const double* invmi, const double* radius,
                                   const uint8_t* in_fov, const double* upsf,
                                   const double* parint, long S, long D, long E,
                                   int mode, double* out) {
    or_psf_radius_integral(radius, upsf, parint, S, D, E, mode, out);
    for (long s = 0; s < S; ++s)
        for (long e = 0; e < E; ++e) {
            double integ = out[s * E + e];
            integ = integ > 1.0 ? 1.0 : integ;
            out[s * E + e] = in_fov[s] ? integ * invmi[s * E + e] : 1.0;
        }
}

/* model_map/model_map.map_correction_factor.cxx:3-53 as coded. */
void or_map_correction_factor(const double* invmi, const double* radius,
                              const uint8_t* in_fov, const double* upsf,
                              const double* parint, long S, long D, long E,
                              double* out) {
    or_map_correction_factor_mode(invmi, radius, in_fov, upsf, parint, S, D, E, 0, out);
}

/* model_map.scale_map_by_solid_angle.cxx:6-26 */
void or_scale_map_by_solid_angle(double* map, long S, long Nh, long Nw, long E,
                                 const double* om) {
    for (long s = 0; s < S; ++s)
        for (long h = 0; h < Nh; ++h)
            for (long w = 0; w < Nw; ++w) {
                double* m = map + (((size_t)s * Nh + h) * Nw + w) * E;
                for (long e = 0; e < E; ++e) m[e] *= om[h * Nw + w];
            }
}

/* model_map.scale_map_by_exposure.cxx:5-28 */
void or_scale_map_by_exposure(double* map, long S, long Nh, long Nw, long E,
                              const double* ex) {
    for (long s = 0; s < S; ++s)
        for (long hw = 0; hw < Nh * Nw; ++hw) {
            double* m = map + ((size_t)s * Nh * Nw + hw) * E;
            for (long e = 0; e < E; ++e) m[e] *= ex[s * E + e];
        }
}

/* model_map.scale_map_by_correction_factor.cxx:4-27 */
void or_scale_map_by_correction_factors(double* map, long S, long Nh, long Nw,
                                        long E, const double* cf) {
    or_scale_map_by_exposure(map, S, Nh, Nw, E, cf);
}

/* model_map/model_map.point_src_model_map_wcs.cxx:6-45: cubature -> solid
 * angle -> boundary radius -> map integral -> exposure -> correction. */
void or_point_src_model_map_wcs(long Nh, long Nw, const or_wcs* wc,
                                const double* src_sph, long S,
                                const double* expo, const double* upsf,
                                const double* parint, long E,
                                const double* rule, long N, double* out,
                                double* radius, uint8_t* in_fov,
                                double* invmi, double* cf) {
    double* pw = (double*)malloc((size_t)Nh * Nw * 4 * N * sizeof(double));
    double* om = (double*)malloc((size_t)Nh * Nw * sizeof(double));
    or_square_ptswts(Nh, Nw, wc, rule, N, pw);
    or_cubature_convolve_map_psf(Nh, Nw, src_sph, S, upsf, E, wc, pw, N,
                                 OR_RULE_LOGSEP, out);
    or_solid_angle(Nh, Nw, wc, om);
    or_scale_map_by_solid_angle(out, S, Nh, Nw, E, om);
    or_psf_boundary_radius(Nh, Nw, src_sph, S, wc, radius, in_fov);
    or_map_integral(out, S, Nh, Nw, E, src_sph, wc, radius, in_fov, invmi);
    or_scale_map_by_exposure(out, S, Nh, Nw, E, expo);
    or_map_correction_factor(invmi, radius, in_fov, upsf, parint, S, OR_ND, E, cf);
    or_scale_map_by_correction_factors(out, S, Nh, Nw, E, cf);
    free(pw);
    free(om);
}

/* ======================================================================
 * Diffuse-source model map (model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97,
 * a function the reference declares but does not compile).  The steps it
 * names, with the choices where its sources are unfinished: the crop box
 * bounds the cube coordinates of the model map's whole border (the
 * reference's fits/fits.read_allsky_cropped.cpp:163-192 takes two corners),
 * unwrapped across the seam of an all-sky CAR cube; the resampled grid is
 * trilerp's corner-aligned grid (math/trilerp.hpp:22-25); the kernel is
 * centred on the model map's reference direction and blends table nodes as
 * mm_cubature/convolve_pixel_psf_logsep.hpp:45-55,72-73 does (the
 * reference's psf/psf.make_kernel.cpp:50, 56 swaps the two weights); the
 * convolution is per energy plane; the convolved grid is sampled
 * bilinearly at the model map's pixel centres, then scaled by solid angle
 * and exposure (diffuse_src_model_map_wcs.cxx:76, 88).
 * ====================================================================== */

/* J2000 galactic frame from the north galactic pole (192.85948, 27.12825)
 * and the galactic longitude of the north celestial pole, 122.93192, as
 * explicit axes: z the pole, x and y in the galactic plane placed so that
 * the celestial pole's projection lies at longitude 122.93192. */
static void gal_axes(double x[3], double y[3], double z[3]) {
    const double ra = 192.85948 * D2R, dec = 27.12825 * D2R, l = 122.93192 * D2R;
    z[0] = cos(dec) * cos(ra); z[1] = cos(dec) * sin(ra); z[2] = sin(dec);
    const double p[3] = {0.0, 0.0, 1.0};
    const double pz = p[0] * z[0] + p[1] * z[1] + p[2] * z[2];
    double u[3] = {p[0] - pz * z[0], p[1] - pz * z[1], p[2] - pz * z[2]};
    const double nu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; ++k) u[k] /= nu;
    const double w[3] = {z[1] * u[2] - z[2] * u[1], z[2] * u[0] - z[0] * u[2],
                         z[0] * u[1] - z[1] * u[0]};
    for (int k = 0; k < 3; ++k) {
        x[k] = cos(l) * u[k] - sin(l) * w[k];
        y[k] = sin(l) * u[k] + cos(l) * w[k];
    }
}

static void vec_sph(const double v[3], double* lon, double* lat) {
    double a = atan2(v[1], v[0]) * R2D;
    if (a < 0.0) a += 360.0;
    if (a >= 360.0) a -= 360.0;
    double z = v[2] > 1.0 ? 1.0 : (v[2] < -1.0 ? -1.0 : v[2]);
    *lon = a;
    *lat = asin(z) * R2D;
}

void or_cel2gal(double ra, double dec, double* l, double* b) {
    double x[3], y[3], z[3], v[3];
    gal_axes(x, y, z);
    or_sph2dir(ra, dec, v);
    const double g[3] = {v[0] * x[0] + v[1] * x[1] + v[2] * x[2],
                         v[0] * y[0] + v[1] * y[1] + v[2] * y[2],
                         v[0] * z[0] + v[1] * z[1] + v[2] * z[2]};
    vec_sph(g, l, b);
}

void or_gal2cel(double l, double b, double* ra, double* dec) {
    double x[3], y[3], z[3], g[3];
    gal_axes(x, y, z);
    or_sph2dir(l, b, g);
    const double v[3] = {g[0] * x[0] + g[1] * y[0] + g[2] * z[0],
                         g[0] * x[1] + g[1] * y[1] + g[2] * z[1],
                         g[0] * x[2] + g[1] * y[2] + g[2] * z[2]};
    vec_sph(v, ra, dec);
}

static void frame_conv(int frame, double lon, double lat, double* olon, double* olat) {
    if (frame == 1) or_cel2gal(lon, lat, olon, olat);
    else if (frame == 2) or_gal2cel(lon, lat, olon, olat);
    else { *olon = lon; *olat = lat; }
}

static int roi2cube(const or_wcs* roi, const or_wcs* cube, int frame, double p1, double p2,
                    double* q1, double* q2) {
    double lon, lat, l2, b2;
    if (or_pix2sph(roi, p1, p2, &lon, &lat) != 0) return -1;
    frame_conv(frame, lon, lat, &l2, &b2);
    or_sph2pix(cube, l2, b2, q1, q2);
    return 0;
}

static long fft_len(long need) {
    for (long n = need < 1 ? 1 : need; n <= 4096; ++n) {
        long m = n;
        while (m % 2 == 0) m /= 2;
        while (m % 3 == 0) m /= 3;
        while (m % 5 == 0) m /= 5;
        if (m == 1) return n;
    }
    return 0;
}

int or_diffuse_geometry(const or_wcs* roi, int roi_gal, const or_wcs* cube, int cube_gal,
                        double target, double margin, or_diffuse_geom* g) {
    memset(g, 0, sizeof(*g));
    g->frame = roi_gal == cube_gal ? 0 : (cube_gal ? 1 : 2);
    const int back = g->frame == 0 ? 0 : 3 - g->frame;
    const long N1 = cube->naxis1, N2 = cube->naxis2, H = roi->naxis1, W = roi->naxis2;
    const double c1 = fabs(cube->cdelt1), c2 = fabs(cube->cdelt2);
    const double turn = cube->proj == OR_PROJ_CAR && cube->crota2 == 0.0 ? 360.0 / c1 : 0.0;
    const int wrap = turn > 0.0 && fabs(N1 * c1 - 360.0) <= 1e-6 * 360.0;
    double x0, y0;
    if (roi2cube(roi, cube, g->frame, 0.5 * (H + 1), 0.5 * (W + 1), &x0, &y0)) return -4;
    double lo1 = x0, hi1 = x0, lo2 = y0, hi2 = y0;
    for (int side = 0; side < 2; ++side) {
        const long n = side == 0 ? H : W;
        for (long k = 0; k <= 2 * (n - 1); ++k) {
            for (int edge = 0; edge < 2; ++edge) {
                double p1, p2, q1, q2;
                if (side == 0) { p1 = 1.0 + 0.5 * k; p2 = edge ? (double)W : 1.0; }
                else { p2 = 1.0 + 0.5 * k; p1 = edge ? (double)H : 1.0; }
                if (roi2cube(roi, cube, g->frame, p1, p2, &q1, &q2)) continue;
                if (turn > 0.0) q1 += turn * nearbyint((x0 - q1) / turn);
                if (q1 < lo1) lo1 = q1;
                if (q1 > hi1) hi1 = q1;
                if (q2 < lo2) lo2 = q2;
                if (q2 > hi2) hi2 = q2;
            }
        }
    }
    int pole = 0;
    for (int k = 0; k < 2 && turn > 0.0; ++k) {
        double lon, lat, p1, p2, q1, q2;
        frame_conv(back, 0.0, k == 0 ? 90.0 : -90.0, &lon, &lat);
        or_sph2pix(roi, lon, lat, &p1, &p2);
        if (p1 >= 0.5 && p1 <= H + 0.5 && p2 >= 0.5 && p2 <= W + 0.5) {
            pole = 1;
            or_sph2pix(cube, 0.0, k == 0 ? 90.0 : -90.0, &q1, &q2);
            if (q2 < lo2) lo2 = q2;
            if (q2 > hi2) hi2 = q2;
        }
    }
    const long mg1 = 1 + (long)ceil(margin / c1), mg2 = 1 + (long)ceil(margin / c2);
    long a1 = (long)floor(lo1) - mg1, b1 = (long)ceil(hi1) + mg1;
    long a2 = (long)floor(lo2) - mg2, b2 = (long)ceil(hi2) + mg2;
    if (pole) {
        a1 = wrap ? (long)floor(x0) - N1 / 2 : 1;
        b1 = wrap ? a1 + N1 - 1 : N1;
    }
    if (wrap) {
        if (b1 - a1 + 1 > N1) b1 = a1 + N1 - 1;
    } else {
        if (a1 < 1) a1 = 1;
        if (b1 > N1) b1 = N1;
    }
    if (a2 < 1) a2 = 1;
    if (b2 > N2) b2 = N2;
    if (b1 - a1 + 1 < 2 || b2 - a2 + 1 < 2) return -1;
    g->a1 = a1; g->a2 = a2; g->n1 = b1 - a1 + 1; g->n2 = b2 - a2 + 1;
    /* sky_geom/sky_geom.hxx:350-372: one pixel size, the target when finer
     * than the cube's coarser axis, naxis = round(naxis * scale) */
    const double pixsz = target < (c1 > c2 ? c1 : c2) ? target : c1;
    g->m1 = lround(g->n1 * (c1 / pixsz));
    g->m2 = lround(g->n2 * (c2 / pixsz));
    if (g->m1 < 2 || g->m2 < 2 || g->m1 > 4096 || g->m2 > 4096) return -2;
    g->step1 = (double)(g->n1 - 1) / (double)(g->m1 - 1);
    g->step2 = (double)(g->n2 - 1) / (double)(g->m2 - 1);
    g->grid = *cube;
    g->grid.naxis1 = g->m1;
    g->grid.naxis2 = g->m2;
    g->grid.crpix1 = (cube->crpix1 - a1) / g->step1 + 1.0;
    g->grid.crpix2 = (cube->crpix2 - a2) / g->step2 + 1.0;
    g->grid.cdelt1 = cube->cdelt1 * g->step1;
    g->grid.cdelt2 = cube->cdelt2 * g->step2;
    or_wcs_setup(&g->grid);
    g->wrap = wrap;
    g->period = wrap ? N1 / g->step1 : 0.0;
    g->vlo1 = wrap ? -INFINITY : (0.5 - a1) / g->step1 + 1.0;
    g->vhi1 = wrap ? INFINITY : (N1 + 0.5 - a1) / g->step1 + 1.0;
    g->vlo2 = (0.5 - a2) / g->step2 + 1.0;
    g->vhi2 = (N2 + 0.5 - a2) / g->step2 + 1.0;
    g->nk = g->m1 < g->m2 ? g->m1 : g->m2;
    if (g->nk % 2 == 0) g->nk -= 1;
    double lon, lat;
    frame_conv(g->frame, roi->crval1, roi->crval2, &lon, &lat);
    or_sph2pix(&g->grid, lon, lat, &g->kc1, &g->kc2);
    if (wrap) g->kc1 += g->period * nearbyint((0.5 * (g->m1 + 1) - g->kc1) / g->period);
    const long c = (g->nk - 1) / 2;
    g->fft_p = fft_len(g->m1 + c);
    g->fft_q = fft_len(g->m2 + c);
    return (g->fft_p > 0 && g->fft_q > 0) ? 0 : -3;
}

/* math/trilerp.hpp:9-56 over the crop of the cube (cube [Ec][N2][N1], FITS
 * order): out [E][m1][m2], out[k][i][j] at crop position (i s1, j s2) and
 * cube plane k se, s = (src - 1) / (dst - 1); crop column i is cube pixel
 * a1 + i along axis 1, taken modulo N1 when the cube wraps. */
void or_diffuse_resample(const float* cube, long N1, long N2, long Ec, const or_diffuse_geom* g,
                         long E, double* out) {
    const double se = E > 1 ? (double)(Ec - 1) / (double)(E - 1) : 0.0;
    const double sh = (double)(g->n1 - 1) / (double)(g->m1 - 1);
    const double sw = (double)(g->n2 - 1) / (double)(g->m2 - 1);
#pragma omp parallel for schedule(static, 1)
    for (long k = 0; k < E; ++k) {
        const double e = k * se;
        const long e0 = (long)e, e1 = e0 + 1 < Ec - 1 ? e0 + 1 : Ec - 1;
        const double fe = e - e0, ce = 1. - fe;
        for (long i = 0; i < g->m1; ++i) {
            const double h = i * sh;
            const long h0 = (long)h, h1 = h0 + 1 < g->n1 - 1 ? h0 + 1 : g->n1 - 1;
            const double fh = h - h0, ch = 1. - fh;
            long x0 = g->a1 + h0 - 1, x1 = g->a1 + h1 - 1;
            if (g->wrap) { x0 = ((x0 % N1) + N1) % N1; x1 = ((x1 % N1) + N1) % N1; }
            for (long j = 0; j < g->m2; ++j) {
                const double w = j * sw;
                const long w0 = (long)w, w1 = w0 + 1 < g->n2 - 1 ? w0 + 1 : g->n2 - 1;
                const double fw = w - w0, cw = 1. - fw;
                const long y0 = g->a2 + w0 - 1, y1 = g->a2 + w1 - 1;
#define CV(ee, xx, yy) ((double)cube[((ee) * N2 + (yy)) * N1 + (xx)])
                const double C00 = ch * CV(e0, x0, y0) + fh * CV(e0, x1, y0);
                const double C01 = ch * CV(e1, x0, y0) + fh * CV(e1, x1, y0);
                const double C10 = ch * CV(e0, x0, y1) + fh * CV(e0, x1, y1);
                const double C11 = ch * CV(e1, x0, y1) + fh * CV(e1, x1, y1);
#undef CV
                const double C0 = cw * C00 + fw * C10;
                const double C1 = cw * C01 + fw * C11;
                out[(k * g->m1 + i) * g->m2 + j] = ce * C0 + fe * C1;
            }
        }
    }
}

/* psf/psf.make_kernel.cpp:8-90: kernel pixel (h, w) at grid pixel
 * (kc1 + h - c, kc2 + w - c), its table value at the separation from the
 * grid direction at (kc1, kc2) (the logsep blend of one unit-weight point,
 * or_pixel_logsep), normalised to sum 1 per energy: out [E][nk][nk]. */
void or_psf_make_kernel(long nk, const or_wcs* grid, double kc1, double kc2,
                        const double* upsf, long E, double* out) {
    const long c = (nk - 1) / 2;
    double d0[3];
    or_pix2dir(grid, kc1, kc2, d0);
    double* val = (double*)malloc((size_t)E * sizeof(double));
    double* tot = (double*)calloc((size_t)E, sizeof(double));
    for (long h = 0; h < nk; ++h) {
        for (long w = 0; w < nk; ++w) {
            double pw[4];
            for (long e = 0; e < E; ++e) val[e] = 0.0;
            /* a kernel pixel outside the grid's projection is 0 */
            if (or_pix2dir(grid, kc1 + (double)(h - c), kc2 + (double)(w - c), pw) == 0) {
                pw[3] = 1.0;
                or_pixel_logsep(val, E, pw, 1, upsf, d0);
            }
            for (long e = 0; e < E; ++e) {
                out[(e * nk + h) * nk + w] = val[e];
                tot[e] += val[e];
            }
        }
    }
    for (long e = 0; e < E; ++e) {
        const double inv = 1.0 / tot[e];
        for (long k = 0; k < nk * nk; ++k) out[e * nk * nk + k] *= inv;
    }
    free(val);
    free(tot);
}

/* Direct linear convolution per plane: out[e][i][j] = sum_{h,w}
 * ker[e][h][w] img[e][i + c - h][j + c - w] (zero outside), i.e. rows and
 * columns c .. c + m - 1 of the full convolution that
 * math/convolve/fft/fft.fftw_convolve_CHW.cpp:7-90 computes by FFT. */
void or_convolve_same(const double* img, long E, long m1, long m2, const double* ker, long nk,
                      long c, double* out) {
#pragma omp parallel for collapse(2) schedule(static, 4)
    for (long e = 0; e < E; ++e) {
        for (long i = 0; i < m1; ++i) {
            const double* K = ker + e * nk * nk;
            const double* I = img + e * m1 * m2;
            for (long j = 0; j < m2; ++j) {
                double s = 0.0;
                for (long h = 0; h < nk; ++h) {
                    const long r = i + c - h;
                    if (r < 0 || r >= m1) continue;
                    for (long w = 0; w < nk; ++w) {
                        const long q = j + c - w;
                        if (q < 0 || q >= m2) continue;
                        s += K[h * nk + w] * I[r * m2 + q];
                    }
                }
                out[(e * m1 + i) * m2 + j] = s;
            }
        }
    }
}

/* The convolved grid [E][m1][m2] sampled bilinearly at the model map's
 * pixel centres (clamped to the grid; 0 outside the cube), times solid
 * angle and exposure: out [E][H][W]. */
void or_diffuse_sample(const double* conv, const or_diffuse_geom* g, const or_wcs* roi,
                       const double* omega, const double* expo, long E, double* out) {
    const long H = roi->naxis1, W = roi->naxis2, m1 = g->m1, m2 = g->m2;
    for (long h = 0; h < H; ++h) {
        for (long w = 0; w < W; ++w) {
            double lon = 0.0, lat = 0.0, l2, b2, v1, v2;
            const int valid = or_pix2sph(roi, h + 1.0, w + 1.0, &lon, &lat) == 0;
            frame_conv(g->frame, lon, lat, &l2, &b2);
            or_sph2pix(&g->grid, l2, b2, &v1, &v2);
            if (g->period > 0.0) v1 += g->period * nearbyint((0.5 * (m1 + 1) - v1) / g->period);
            const int inside = valid && v1 >= g->vlo1 && v1 <= g->vhi1 && v2 >= g->vlo2 && v2 <= g->vhi2;
            double x = v1 - 1.0, y = v2 - 1.0;
            x = x < 0.0 ? 0.0 : (x > m1 - 1 ? (double)(m1 - 1) : x);
            y = y < 0.0 ? 0.0 : (y > m2 - 1 ? (double)(m2 - 1) : y);
            long i0 = (long)x, j0 = (long)y;
            if (i0 > m1 - 2) i0 = m1 - 2;
            if (j0 > m2 - 2) j0 = m2 - 2;
            const double fx = x - i0, fy = y - j0;
            for (long e = 0; e < E; ++e) {
                const double* c = conv + (e * m1 + i0) * m2 + j0;
                const double v = (1.0 - fx) * (1.0 - fy) * c[0] + (1.0 - fx) * fy * c[1] +
                                 fx * (1.0 - fy) * c[m2] + fx * fy * c[m2 + 1];
                out[(e * H + h) * W + w] = inside ? v * omega[h * W + w] * expo[e] : 0.0;
            }
        }
    }
}

int or_diffuse_src_model_map_wcs(const or_wcs* roi, int roi_gal, const or_wcs* cube,
                                 int cube_gal, const float* data, long Ec, const double* upsf,
                                 const double* expo, long E, double target, double margin,
                                 double* out) {
    or_diffuse_geom g;
    const int rc = or_diffuse_geometry(roi, roi_gal, cube, cube_gal, target, margin, &g);
    if (rc) return rc;
    const size_t plane = (size_t)g.m1 * g.m2;
    double* img = (double*)malloc((size_t)E * plane * sizeof(double));
    double* ker = (double*)malloc((size_t)E * g.nk * g.nk * sizeof(double));
    double* conv = (double*)malloc((size_t)E * plane * sizeof(double));
    double* om = (double*)malloc((size_t)roi->naxis1 * roi->naxis2 * sizeof(double));
    or_diffuse_resample(data, cube->naxis1, cube->naxis2, Ec, &g, E, img);
    or_psf_make_kernel(g.nk, &g.grid, g.kc1, g.kc2, upsf, E, ker);
    or_convolve_same(img, E, g.m1, g.m2, ker, g.nk, (g.nk - 1) / 2, conv);
    or_solid_angle(roi->naxis1, roi->naxis2, roi, om);
    or_diffuse_sample(conv, &g, roi, om, expo, E, out);
    free(img);
    free(ker);
    free(conv);
    free(om);
    return 0;
}
