/*
 * xts_oracle.h -- FP64 CPU restatement of the xtsrcmaps point-source model-map
 * path: (a) exposure contraction, (b) mean-PSF build, (c) pixel cubature and
 * the model-map scaling passes, and the diffuse-source map.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; nothing in xtsrcmaps_amd/ links it.
 *
 * Every function cites the reference file:line range it restates.  Arrays are
 * row-major, sizes in elements.  All arithmetic is IEEE double.
 *
 * Pins (tests/test_oracle_golden.py, tests/test_oracle_invariants.py): the
 * reference's fixtures pin the separations, prepare_grid, normalize_rpsf,
 * king, LTF, costhetas, aeff_value, the exposure contractions, the mean-PSF
 * peak, the pixel cubature maps (logsep, searchsep, three rules) and the
 * FermiTools in-FOV flags and radii; closed forms and invariants pin
 * ang2pix, CAR, the WCS round trip and CROTA2, the Gauss-Legendre rule, the
 * planar King integral, the cumulative radial integral, constant-LUT
 * cubature, logsep == searchsep inside the grid, a brute-force pixel
 * integral, and map_integral / map_correction_factor through the
 * constant-LUT pipeline identity at a 1-degree radius.
 *
 * Pinned by mathematics the tests compute themselves
 * (tests/test_oracle_pins.py): or_solid_angle as the spherical area of the
 * neighbouring-centre quad (oblique rotated CAR, sheared AIT, and AIT's
 * equal area), where a corner quad, one triangle or a shifted quad misses;
 * or_map_correction_factor's node index and segment for every radius
 * (including the index-1 branch below 1 degree) against the exact integral
 * of a piecewise-linear integrand; or_pixel_logsep outside [1e-4, 70]
 * degrees against the edge segments' lines; or_psf_radius_integral's
 * exact-radius variant against FermiTools' ft_mm_upsf_radius_integ.bin.
 *
 * The diffuse-source map (tests/test_oracle_diffuse.py): or_cel2gal /
 * or_gal2cel against the galactic centre's published J2000 position and the
 * poles; or_diffuse_geometry by the crop holding every map pixel centre and
 * the grid's pixel-to-cube map; or_diffuse_resample exact on trilinear
 * cubes and equal across the longitude seam to a rolled cube;
 * or_psf_make_kernel against a table linear in separation (closed form);
 * or_convolve_same against scipy.signal.convolve2d; or_diffuse_sample
 * exact on a linear grid.
 *
 * Parity unpinned: none.
 */
#ifndef XTS_ORACLE_H
#define XTS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_ND = 401 };          /* psf/psf.hxx:20 sep_arr_len */
enum { OR_PROJ_CAR = 0, OR_PROJ_AIT = 1 };
enum { OR_RULE_LOGSEP = 0, OR_RULE_SEARCHSEP = 1 };

typedef struct {
    long   naxis1, naxis2;
    double crpix1, crpix2, crval1, crval2, cdelt1, cdelt2, crota2;
    int    proj;
    /* derived by or_wcs_setup */
    double m11, m12, m21, m22;     /* pixel offset -> intermediate (deg) */
    double i11, i12, i21, i22;     /* inverse of the above */
    double alphap, deltap, phip;   /* Euler angles (deg) */
} or_wcs;

/* separations and constants (psf/psf.hxx:24-43, misc/misc.hxx:11-39) */
void   or_separations(double* out401);
double or_sep_step(void);
int    or_fast_separation_lower_index(double v);

/* lerp_pars (math/bilerp.cxx:6-27) */
int or_lerp_pars(const double* grid, long n, double v, int has_zlb, double zlb,
                 double* w, double* cw, long* idx);

/* IRF math (irf/irf.hxx:22-77, irf/irf.normalize_rpsf.cpp:11-61,
 * math/gauss_legendre.hxx:12-84, irf/irf.cxx:70-109, irf.prepare_grid.cpp) */
double or_evaluate_king(double sep_rad, double sf, const double* pars6);
double or_psf_base_integral(double radius_deg, double sf, const double* pars6);
void   or_legendre_rw(int n, double eps, double* x, double* w);
void   or_normalize_rpsf(double* params, const double* logEs, long Mc, long Me,
                         float scale0, float scale1, float scale_index);
void   or_prepare_grid(const float* e_lo, const float* e_hi, long Me0,
                       const float* c_lo, const float* c_hi, long Mc0,
                       const float* raw, long Ng,
                       double* cosths, double* logEs, double* params,
                       double* min_costheta);
void   or_ltf(const double* logEs, long E, const float* p0, const float* p1,
              double* out2E);
int    or_aeff_value(const double* costh, long C, const double* logE, long E,
                     const double* pc, long Mc, const double* pe, long Me,
                     const double* params, double min_costheta, double* out);

/* (a) exposure (exposure/exposure.cxx:15-173, exposure.contract.cpp:13-42,
 * healpix/healpix.cxx:182-229) */
void     or_costhetas(long nbins, double cosmin, int thetabin, double* out);
uint64_t or_ang2pix_nest(double theta, double phi, int64_t nside);
void     or_src_cosbins(const double* src_sph, long S, const double* cube,
                        int64_t nside, long C, double* out);
void     or_contract(const double* A, const double* B, long M, long K, long N,
                     double* out);
void     or_exposure(const double* ecb, const double* wecb, const double* af,
                     const double* ab, const double* ltf, long S, long C,
                     long E, double* out, double* lef, double* lwf,
                     double* leb, double* lwb);

/* (b) mean PSF (psf/psf.*.cxx) */
void or_king(const double* params, long Mc, long Me, double* out);
int  or_bilerp(const double* costh, long C, const double* logE, long E,
               const double* pc, long Mc, const double* pe, long Me,
               const double* kings, double* out);
void or_corrected_exposure_psf(const double* obs_psf, const double* aeff,
                               const double* ecb, const double* wecb,
                               const double* ltf, long S, long C, long D,
                               long E, double* out);
void or_mean_psf(const double* front, const double* back, const double* expo,
                 long S, long D, long E, double* out);
void or_partial_total_integral(const double* upsf, long S, long D, long E,
                               double* parint, double* totint);
void or_normalize(double* upsf, const double* totint, long S, long D, long E);
int  or_psf_lookup_table_and_partial_integrals(
        const double* params_f, const double* pc_f, long Mc_f,
        const double* pe_f, long Me_f,
        const double* params_b, const double* pc_b, long Mc_b,
        const double* pe_b, long Me_b,
        const double* costh, long C, const double* logE, long E,
        const double* af, const double* ab, const double* ecb,
        const double* wecb, const double* ltf, const double* expo, long S,
        double* upsf, double* parint);

/* WCS (sky_geom/sky_geom.hxx:79-310; wcslib semantics, Calabretta & Greisen
 * 2002 for CAR and AIT) */
void or_wcs_setup(or_wcs* w);
int  or_pix2sph(const or_wcs* w, double p1, double p2, double* lon, double* lat);
int  or_sph2pix(const or_wcs* w, double lon, double lat, double* p1, double* p2);
void or_sph2dir(double lon, double lat, double* dir3);
int  or_pix2dir(const or_wcs* w, double p1, double p2, double* dir3);
double or_dir_diff(const double* a, const double* b);

/* (c) model map (model_map/ and model_map/mm_cubature/) */
void or_square_ptswts(long Nh, long Nw, const or_wcs* w, const double* rule,
                      long N, double* out);
void or_naive_ptswts(long Nh, long Nw, const or_wcs* w, long npix, double* out);
void or_pixel_logsep(double* map, long E, const double* ptswts, long N,
                     const double* lut, const double* src3);
void or_pixel_searchsep(double* map, long E, const double* ptswts, long N,
                        const double* lut, const double* src3);
void or_cubature_convolve_map_psf(long Nh, long Nw, const double* src_sph,
                                  long S, const double* lut, long E,
                                  const or_wcs* w, const double* ptswts,
                                  long N, int mode, double* out);
void or_solid_angle(long Nh, long Nw, const or_wcs* w, double* out);
void or_psf_boundary_radius(long Nh, long Nw, const double* src_sph, long S,
                            const or_wcs* w, double* radius, uint8_t* in_fov);
void or_map_integral(const double* map, long S, long Nh, long Nw, long E,
                     const double* src_sph, const or_wcs* w,
                     const double* radius, const uint8_t* in_fov, double* out);
void or_psf_radius_integral(const double* radius, const double* upsf,
                            const double* parint, long S, long D, long E,
                            int mode, double* out);
void or_map_correction_factor_mode(const double* invmi, const double* radius,
                                   const uint8_t* in_fov, const double* upsf,
                                   const double* parint, long S, long D, long E,
                                   int mode, double* out);
void or_map_correction_factor(const double* inv_mapinteg, const double* radius,
                              const uint8_t* in_fov, const double* upsf,
                              const double* parint, long S, long D, long E,
                              double* out);
void or_scale_map_by_solid_angle(double* map, long S, long Nh, long Nw, long E,
                                 const double* omega);
void or_scale_map_by_exposure(double* map, long S, long Nh, long Nw, long E,
                              const double* expo);
void or_scale_map_by_correction_factors(double* map, long S, long Nh, long Nw,
                                        long E, const double* cf);
void or_point_src_model_map_wcs(long Nh, long Nw, const or_wcs* w,
                                const double* src_sph, long S,
                                const double* expo, const double* upsf,
                                const double* parint, long E,
                                const double* rule, long N, double* out,
                                double* radius, uint8_t* in_fov,
                                double* inv_mapinteg, double* cf);
int  or_num_threads(void);

/* diffuse-source model map (model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97
 * and the steps it names; see the note above or_cel2gal in xts_oracle.c) */
typedef struct {
    long   a1, a2, n1, n2;     /* crop start (1-based cube pixels) and size */
    long   m1, m2;             /* resampled grid */
    double step1, step2;       /* cube pixels per grid pixel */
    or_wcs grid;               /* WCS of the resampled grid */
    int    wrap, frame;        /* longitude wrap; 0 same frame, 1 cel->gal, 2 gal->cel */
    long   nk;                 /* kernel size */
    double kc1, kc2;           /* kernel centre on the grid */
    long   fft_p, fft_q;       /* transform lengths */
    double period;             /* grid pixels per turn (wrap) */
    double vlo1, vhi1, vlo2, vhi2;  /* grid coordinates inside the cube */
} or_diffuse_geom;
void or_cel2gal(double ra, double dec, double* l, double* b);
void or_gal2cel(double l, double b, double* ra, double* dec);
int  or_diffuse_geometry(const or_wcs* roi, int roi_gal, const or_wcs* cube, int cube_gal,
                         double target, double margin, or_diffuse_geom* g);
void or_diffuse_resample(const float* cube, long N1, long N2, long Ec,
                         const or_diffuse_geom* g, long E, double* out);
void or_psf_make_kernel(long nk, const or_wcs* grid, double kc1, double kc2,
                        const double* upsf, long E, double* out);
void or_convolve_same(const double* img, long E, long m1, long m2, const double* ker,
                      long nk, long c, double* out);
void or_diffuse_sample(const double* conv, const or_diffuse_geom* g, const or_wcs* roi,
                       const double* omega, const double* expo, long E, double* out);
int  or_diffuse_src_model_map_wcs(const or_wcs* roi, int roi_gal, const or_wcs* cube,
                                  int cube_gal, const float* data, long Ec,
                                  const double* upsf, const double* expo, long E,
                                  double target, double margin, double* out);

#ifdef __cplusplus
}
#endif
#endif
