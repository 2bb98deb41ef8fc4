/*
 * xtsrcmaps_amd.h -- C ABI of the MI355X (gfx950) point-source model-map path.
 *
 * Each entry point replaces one function of the reference xtsrcmaps C++ API
 * (cited as file:line in Areustle/xtsrcmaps).  Arrays are row-major; sizes are
 * element counts.  Host-pointer entry points copy in, compute on the device
 * and copy out before returning.  `_dev` entry points take device pointers,
 * enqueue on the context's stream and return without synchronising.
 *
 * Every function returns XTS_OK (0) or an XTS_ERR_* code; xts_last_error()
 * gives the message of the last failure on the calling thread.  No C++
 * exception crosses this boundary.  A context is used by one host thread at
 * a time.
 */
#ifndef XTSRCMAPS_AMD_H
#define XTSRCMAPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XTS_ABI_VERSION 1

/* status codes */
#define XTS_OK 0
#define XTS_ERR_INVALID 1 /* null pointer, bad size or flag              */
#define XTS_ERR_DEVICE 2  /* HIP runtime failure / no device             */
#define XTS_ERR_ALLOC 3   /* device allocation failed                    */
#define XTS_ERR_RANGE 4   /* a sample lies outside an IRF grid           */

/* projections */
#define XTS_PROJ_CAR 0
#define XTS_PROJ_AIT 1

/* model-map flags */
#define XTS_LAYOUT_SHWE 0x0   /* [S, H, W, E]: the reference's layout          */
#define XTS_LAYOUT_SEHW 0x1   /* [S, E, H, W]: energy planes, coalesced stores */
#define XTS_OUT_F64 0x0       /* double output                                 */
#define XTS_OUT_F32 0x2       /* float output                                  */
#define XTS_COMPUTE_F32 0x0   /* FP32 cubature (FP64 geometry and scaling)     */
#define XTS_COMPUTE_F64 0x4   /* FP64 cubature throughout                      */
/* Correction factor with the PSF integrated to exactly the boundary radius
 * from the node below it, the FermiTools-consistent variant; without it, as
 * model_map.map_correction_factor.cxx:24-50 codes it (node index 1 below one
 * degree, the whole next segment). */
#define XTS_CF_EXACT_RADIUS 0x8
/* Far pixels whose rule points span one table node take the learned
 * surrogate (xts_ctx_set_surrogate) instead of the point loop; FP32 cubature
 * only.  No counterpart in the reference. */
#define XTS_SURROGATE 0x10
/* With XTS_SURROGATE: the same pixels without the network, their absolute
 * first moment taken as |A| (the value at the points' mean separation): the
 * method without the network, for comparison. */
#define XTS_SURROGATE_NO_NET 0x20
/* With XTS_SURROGATE: the network's hidden layers on the FP8 (OCP e4m3)
 * matrix cores, one weight scale per layer and activations scaled per wave
 * and layer, FP32 accumulation (the first layer stays BF16 high and low
 * parts).  No counterpart in the reference. */
#define XTS_SURROGATE_FP8 0x40

/* contraction precision for the exposure and mean-PSF builds */
#define XTS_PREC_FP64 0       /* FP64 FMA                                      */
#define XTS_PREC_BF16X3 1     /* BF16 hi/mid/lo split on MFMA, FP32 accumulate */

/* scale kinds for xts_scale_map */
#define XTS_SCALE_PIXEL 0     /* factor [H, W]  (solid angle)                  */
#define XTS_SCALE_SRC_ENERGY 1/* factor [S, E]  (exposure, correction)         */

typedef struct xts_ctx xts_ctx;

/* WCS of a model-map image.  Same fields as the reference's Fermi::WcsConfig
 * (sky_geom/sky_geom.hxx:17-25); CROTA2 in degrees, CDELT in degrees. */
typedef struct {
    int64_t naxis1, naxis2;
    double crpix1, crpix2;
    double crval1, crval2;
    double cdelt1, cdelt2;
    double crota2;
    int32_t proj;
    int32_t is_galactic;
} xts_wcs;

/* One section (front or back) of the King PSF IRF grid after
 * normalisation: Irf::IrfData3 (irf/irf_types.hxx:9-14). */
typedef struct {
    const double* params; /* [Mc, Me, 6] NCORE NTAIL SCORE STAIL GCORE GTAIL */
    const double* cosths; /* [Mc] */
    const double* logEs;  /* [Me] */
    int64_t Mc, Me;
} xts_psf_grid;

/* Effective-area grid: Irf::IrfData3 with one parameter per node. */
typedef struct {
    const double* params; /* [Mc, Me] m^2 */
    const double* cosths; /* [Mc] */
    const double* logEs;  /* [Me] */
    int64_t Mc, Me;
    double min_costheta;
} xts_aeff_grid;

/* ------------------------------------------------------------------ context */
int xts_ctx_create(int device, xts_ctx** out);
int xts_ctx_destroy(xts_ctx* ctx);
/* Use a caller-owned hipStream_t (NULL restores the context's own stream). */
int xts_ctx_set_stream(xts_ctx* ctx, void* hip_stream);
int xts_ctx_synchronize(xts_ctx* ctx);
const char* xts_last_error(void);
int xts_abi_version(void);
/* Instrumentation (no reference counterpart): when enabled, kernel launches
 * of the named model-map stages ("cubature": the row-weight pass, "near":
 * the near-pixel rows, "factor", "emit", "wide" and "near_emit": the map
 * values of the far, wide and near pixels; the row weights of the pipelined
 * map and the surrogate's network pass there run on its second stream and
 * are not timed) are bracketed by events on the
 * context stream; xts_ctx_kernel_times synchronises and returns (and
 * clears) the summed milliseconds and the launch count of one stage. */
int xts_ctx_enable_timing(xts_ctx* ctx, int enable);
/* Launch-shape override for the model map (no reference counterpart):
 * "tiles_per_block" (16x16-pixel tiles per block of the row-weight pass;
 * 0 restores the default choice); "geometry_cache" (1, the default: the
 * pixel geometry and solid angles of the last image are kept between model
 * maps of that image; 0: recomputed by every call); "pipeline_batches" (the
 * FP32 [S, E, H, W] model map in this many source batches, each batch's row
 * weights computed by the same launch as the previous batch's map values;
 * 1: the passes over all sources at once; 0: the default); "surrogate_beside"
 * (the pipelined map's network pass on the second stream as this many
 * resident blocks per CU, default 4; 0: on the context stream).  Results do
 * not depend on any of them.  "surrogate_bf16_lo" (0, the default: the BF16
 * network's hidden weights as plain BF16; 1: as BF16 high and low parts,
 * FP32-level weights) changes the surrogate's values by rounding (about
 * 3e-6 relative on the 2,000-source 400x400 map).  Setting a key cancels a
 * pending first half of a split model map. */
int xts_ctx_set_tuning(xts_ctx* ctx, const char* key, int64_t value);
int xts_ctx_kernel_times(xts_ctx* ctx, const char* name, double* total_ms,
                         int64_t* count);
/* The shape the last batched model map ran with: "batch_sources" (0: the
 * default of about six batches), "rowweight_blocks" (0: the full grid),
 * "batches"; and the counts "graph_captures", "graph_replays" of the
 * context's batched calls.  No reference counterpart. */
int xts_ctx_get_tuning(xts_ctx* ctx, const char* key, int64_t* value);

/* ------------------------------------------------------------- (a) exposure */
/* Exposure::costhetas (exposure/exposure.cxx:29-40). */
int xts_costhetas(int64_t nbins, double cosmin, int32_t thetabin, double* out);

/* Irf::livetime_efficiency_factors (irf/irf.cxx:70-109): p0 -> row 1,
 * p1 -> row 0, breaks formed in float as the IRF columns are float. */
int xts_livetime_efficiency_factors(const double* logEs, int64_t E,
                                    const float* p0, const float* p1,
                                    double* ltf /* [2, E] */);

/* Irf::aeff_value (irf/irf.aeff_value.cxx:11-55). */
int xts_aeff_value(xts_ctx* ctx, const double* costh, int64_t C,
                   const double* logE, int64_t E, const xts_aeff_grid* grid,
                   double* out /* [C, E] */);

/* Device variant: grid arrays, costh, logE and out are device pointers.
 * Samples outside the grid set *range_error (a device int32 the caller
 * clears; NULL: not reported) and give 0. */
int xts_aeff_value_dev(xts_ctx* ctx, const double* costh, int64_t C,
                       const double* logE, int64_t E, const xts_aeff_grid* grid,
                       double* out, int32_t* range_error);

/* Exposure::src_cosbins (exposure/exposure.cxx:42-63) with
 * Healpix::ang2pix NESTED (healpix/healpix.cxx:182-229). */
int xts_src_cosbins(xts_ctx* ctx, const double* src_sph /* [S, 2] deg */,
                    int64_t S, const double* cube /* [12 nside^2, C] */,
                    int64_t nside, int64_t C, double* out /* [S, C] */);
int xts_src_cosbins_dev(xts_ctx* ctx, const double* src_sph, int64_t S,
                        const double* cube, int64_t nside, int64_t C,
                        double* out);

/* Exposure::contract (exposure/exposure.contract.cpp:13-42). */
int xts_contract(xts_ctx* ctx, const double* A /* [M, K] */,
                 const double* B /* [K, N] */, int64_t M, int64_t K,
                 int64_t N, double* out /* [M, N] */, int32_t precision);

/* Exposure::exposure (exposure/exposure.cxx:69-173): four [S,C].[C,E]
 * contractions scaled by the front LTF rows and summed. */
int xts_exposure(xts_ctx* ctx, const double* ecb, const double* wecb,
                 const double* aeff_front, const double* aeff_back,
                 const double* front_ltf, int64_t S, int64_t C, int64_t E,
                 double* exposure /* [S, E] */, int32_t precision);
int xts_exposure_dev(xts_ctx* ctx, const double* ecb, const double* wecb,
                     const double* aeff_front, const double* aeff_back,
                     const double* front_ltf, int64_t S, int64_t C, int64_t E,
                     double* exposure, int32_t precision);

/* ------------------------------------------------------------- (b) mean PSF */
/* irf_private::normalize_rpsf (irf/irf.normalize_rpsf.cpp:11-61), in place. */
int xts_normalize_rpsf(xts_ctx* ctx, double* params /* [Mc, Me, 6] */,
                       const double* logEs, int64_t Mc, int64_t Me,
                       float scale0, float scale1, float scale_index);

/* Psf::king (psf/psf.king.cxx:9-43): [401, Mc, Me]. */
int xts_king(xts_ctx* ctx, const xts_psf_grid* grid, double* out);

/* Psf::psf_lookup_table_and_partial_integrals
 * (psf/psf.psf_lookup_table_and_partial_integrals.cxx:8-57): King on the
 * 401 separations, bilinear to (cos theta, log E), livetime x Aeff weighted
 * contraction, mean over exposure, radial integral and normalisation.
 * upsf and parint are [S, 401, E]. */
int xts_psf_lookup_table_and_partial_integrals(
    xts_ctx* ctx, const xts_psf_grid* front, const xts_psf_grid* back,
    const double* costh, int64_t C, const double* logE, int64_t E,
    const double* aeff_front, const double* aeff_back, const double* ecb,
    const double* wecb, const double* front_ltf, const double* exposure,
    int64_t S, double* upsf, double* parint, int32_t precision);
/* Device variant; grid arrays are device pointers too. */
int xts_psf_lookup_table_and_partial_integrals_dev(
    xts_ctx* ctx, const xts_psf_grid* front, const xts_psf_grid* back,
    const double* costh, int64_t C, const double* logE, int64_t E,
    const double* aeff_front, const double* aeff_back, const double* ecb,
    const double* wecb, const double* front_ltf, const double* exposure,
    int64_t S, double* upsf, double* parint, int32_t precision);

/* Psf::bilerp (psf/psf.bilerp.cxx:6-52): King table kings [401, Mc, Me]
 * on the grid's (cos theta, log E) nodes -> [C, 401, E] at the observation
 * points; grid->params is not read.  XTS_ERR_RANGE if a point lies outside
 * the grid. */
int xts_bilerp(xts_ctx* ctx, const double* costh, int64_t C, const double* logE,
               int64_t E, const xts_psf_grid* grid, const double* kings,
               double* out);

/* Psf::corrected_exposure_psf (psf/psf.corrected_exposure_psf.cxx:12-120):
 * (ECB . (psf x aeff)) LTF0 + (WECB . (psf x aeff)) LTF1 -> [S, 401, E]. */
int xts_corrected_exposure_psf(xts_ctx* ctx, const double* obs_psf /* [C, 401, E] */,
                               const double* obs_aeff /* [C, E] */,
                               const double* ecb, const double* wecb,
                               const double* front_ltf /* [2, E] */, int64_t S,
                               int64_t C, int64_t E, double* out);

/* Psf::mean_psf (psf/psf.mean_psf.cxx:6-33): (front + back) / exposure. */
int xts_mean_psf(xts_ctx* ctx, const double* front, const double* back,
                 const double* exposure /* [S, E] */, int64_t S, int64_t E,
                 double* out /* [S, 401, E] */);

/* Psf::normalize (psf/psf.normalize.cxx:7-24): upsf /= total, in place. */
int xts_normalize(xts_ctx* ctx, double* upsf /* [S, 401, E] */,
                  const double* total /* [S, E] */, int64_t S, int64_t E);

/* Psf::partial_total_integral + Psf::normalize
 * (psf/psf.partial_total_integral.cxx:8-52, psf/psf.normalize.cxx:7-24):
 * upsf is normalised in place, parint [S, 401, E] and total [S, E] out. */
int xts_partial_total_integral(xts_ctx* ctx, double* upsf, int64_t S,
                               int64_t E, double* parint, double* total);

/* ------------------------------------------------------------ (c) model map */
/* ModelMap::point_src_model_map_wcs
 * (model_map/model_map.point_src_model_map_wcs.cxx:6-45): cubature of the
 * PSF table over every pixel, x solid angle, x exposure, x correction factor.
 * rule is [N, 3] (u, v, w) on [-1, 1]^2 with sum w = 4; NULL selects the
 * reference's production rule (Festa-Sommariva degree 7, 12 points).
 * out is laid out and typed by flags (XTS_LAYOUT_*, XTS_OUT_*,
 * XTS_COMPUTE_*). */
int xts_point_src_model_map_wcs(xts_ctx* ctx, const xts_wcs* wcs,
                                const double* src_sph, int64_t S,
                                const double* exposure, const double* upsf,
                                const double* parint, int64_t E,
                                const double* rule, int64_t N, int32_t flags,
                                void* out);
int xts_point_src_model_map_wcs_dev(xts_ctx* ctx, const xts_wcs* wcs,
                                    const double* src_sph, int64_t S,
                                    const double* exposure, const double* upsf,
                                    const double* parint, int64_t E,
                                    const double* rule, int64_t N,
                                    int32_t flags, void* out);
/* xts_point_src_model_map_wcs_dev for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * result; it calls no hipMalloc, hipFree or allocating routine: every buffer
 * besides the inputs and out comes from a fixed 256 MiB device scratch that
 * each device holds from the library's load until the process ends.  A call
 * needing more (about 36 bytes per source and pixel, 400 KB per source of
 * near-pixel pool, plus the image's geometry and the PSF table copy: the
 * 100-source 200x200 37-energy map fits) fails with XTS_ERR_INVALID before
 * any launch.  Calls that overlap on one device (two contexts, two streams)
 * share the scratch and race; the scratch keeps the last call's contents,
 * and everything that must start at zero is cleared by the call's kernels. */
int xts_point_src_model_map_wcs_dev_noalloc(xts_ctx* ctx, const xts_wcs* wcs,
                                            const double* src_sph, int64_t S,
                                            const double* exposure, const double* upsf,
                                            const double* parint, int64_t E,
                                            const double* rule, int64_t N,
                                            int32_t flags, void* out);
/* xts_point_src_model_map_wcs_dev in two calls on the context's stream.
 * The first needs only the image, the sources and the rule: pixel geometry,
 * boundary radii, the per-pixel row weights of the PSF table and the
 * near-pixel rows.  The second (same arguments plus the PSF table, exposure
 * and output) converts the table, forms the map integral and correction
 * factor, and writes the map.  A caller that computes the PSF on another
 * stream (another context) can run it beside the first half, then make this
 * context's stream wait for it before the second.  Between the two calls the
 * context runs no other model map.  The result equals the one-call entry
 * point's.  No counterpart in the reference, whose model map is one function
 * (model_map/model_map.point_src_model_map_wcs.cxx:6-45). */
int xts_point_src_model_map_rows_dev(xts_ctx* ctx, const xts_wcs* wcs,
                                     const double* src_sph, int64_t S, int64_t E,
                                     const double* rule, int64_t N, int32_t flags);
int xts_point_src_model_map_finish_dev(xts_ctx* ctx, const xts_wcs* wcs,
                                       const double* src_sph, int64_t S,
                                       const double* exposure, const double* upsf,
                                       const double* parint, int64_t E,
                                       const double* rule, int64_t N,
                                       int32_t flags, void* out);

/* ------------------------------------------------- batched model maps */
/* Many sources over several images in one call (no counterpart in the
 * reference, which maps one image's sources in one dense loop,
 * mm_cubature/cubature.implementation.cxx:33-53).  The sources of image i
 * are [src_begin[i], src_begin[i+1]) of the source arrays (src_begin[0] = 0,
 * n_images + 1 entries, host memory); wcs [n_images] host.  Each source's
 * values are written for the pixels of its footprint box only, as an FP32
 * [E][h1 - h0][w1 - w0] block at `offset` elements of the output. */
typedef struct {
    int64_t h0, h1, w0, w1; /* pixel indices [h0, h1) x [w0, w1) of its image */
    int64_t offset;         /* element offset of its block in the output */
} xts_footprint;

/* Footprint boxes: per source the largest over energies of the separation
 * at which the normalised cumulative PSF integral (parint, device [S, 401,
 * E]) reaches `containment` (e.g. 0.995; psf.partial_total_integral.cxx:
 * 23-48), plus the pixel half-diagonal (radius [S] deg, host, may be NULL),
 * and the pixel box of that disk about the source; consecutive offsets;
 * *total = the output's element count.  src_sph device [S, 2].  Host
 * results: synchronises the context stream. */
int xts_model_map_footprints(xts_ctx* ctx, int64_t n_images, const xts_wcs* wcs,
                             const int64_t* src_begin, const double* src_sph,
                             const double* parint, int64_t E, double containment,
                             xts_footprint* boxes, double* radius, int64_t* total);

/* The flux-weighted summed model of S point sources on one image: out
 * device float [E, H, W] = sum_s flux[s] x (source s's map of
 * xts_point_src_model_map_wcs_dev).  No counterpart in the reference, whose
 * callers receive one map per source (model_map.point_src_model_map_wcs.cxx
 * via ModelMap::compute_srcmaps) and sum them themselves; it is for sky
 * models whose per-source maps do not fit in memory (10,000 sources x 100
 * energies x 1000 x 1000 pixels is 4 TB).  flux device double [S]; the other
 * inputs as for xts_point_src_model_map_wcs_dev; FP32 cubature; flags may
 * hold XTS_SURROGATE (with XTS_SURROGATE_NO_NET) and XTS_CF_EXACT_RADIUS.
 * The sources run in batches (tuning key "sum_batch_sources"; default: the
 * batch's working buffers within about 12 GB): per batch the row weights,
 * near pixels and factors, then every pixel's sum over the batch's sources
 * accumulated in registers and added to out. */
int xts_point_src_summed_model_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph,
                                   int64_t S, const double* flux, const double* expo,
                                   const double* upsf, const double* parint, int64_t E,
                                   const double* rule, int64_t N, int32_t flags, float* out);

/* The batched model map: for each source, xts_point_src_model_map_wcs_dev's
 * values (its own image, exposure, PSF table, correction factor) inside its
 * box.  Device pointers as for that entry point; boxes host [S] (from
 * xts_model_map_footprints); out device float [out_elems]; flags must hold
 * XTS_LAYOUT_SEHW | XTS_OUT_F32 and FP32 cubature (XTS_SURROGATE and
 * XTS_CF_EXACT_RADIUS apply).  The work runs as a cost-ordered list of
 * source batches alternating between the context's stream and a second
 * internal stream, pipelined: each batch's row weights and factors wait for
 * the previous batch's, its emit for the previous emit, so the row weights
 * of one batch run beside the emit of the previous one; it is ordered after
 * the work already on the context stream and the context stream waits for
 * all of it.  Tuning keys: "batch_sources",
 * "rowweight_blocks", "autotune" (time candidate shapes on the first call
 * of a workload shape, keep the fastest), "graph" (capture the launches in a
 * HIP graph on the first call with given arguments, replay afterwards). */
int xts_point_src_model_maps_batched_dev(xts_ctx* ctx, int64_t n_images, const xts_wcs* wcs,
                                         const int64_t* src_begin, const double* src_sph,
                                         const double* exposure, const double* upsf,
                                         const double* parint, int64_t E, const double* rule,
                                         int64_t N, int32_t flags, const xts_footprint* boxes,
                                         float* out, int64_t out_elems);

/* cubature_convolve_map_psf (mm_cubature/cubature.implementation.cxx:21-55):
 * the raw cubature only. */
int xts_cubature_convolve_map_psf(xts_ctx* ctx, const xts_wcs* wcs,
                                  const double* src_sph, int64_t S,
                                  const double* upsf, int64_t E,
                                  const double* rule, int64_t N,
                                  int32_t flags, void* out);

/* Math::solid_angle (math/solid_angle.hpp:42-79): [H, W]. */
int xts_solid_angle(xts_ctx* ctx, const xts_wcs* wcs, double* out);

/* ModelMap::psf_boundary_radius
 * (model_map/model_map.psf_boundary_radius.cxx:9-62). */
int xts_psf_boundary_radius(xts_ctx* ctx, const xts_wcs* wcs,
                            const double* src_sph, int64_t S, double* radius,
                            uint8_t* in_fov);

/* ModelMap::map_integral (model_map/model_map.map_integral.cxx:4-48):
 * map [S, H, W, E] double -> inverse integral [S, E]. */
int xts_map_integral(xts_ctx* ctx, const xts_wcs* wcs, const double* map,
                     const double* src_sph, int64_t S, int64_t E,
                     const double* radius, const uint8_t* in_fov,
                     double* out);

/* ModelMap::map_correction_factor
 * (model_map/model_map.map_correction_factor.cxx:3-53). */
int xts_map_correction_factor(xts_ctx* ctx, const double* inv_mapinteg,
                              const double* radius, const uint8_t* in_fov,
                              const double* upsf, const double* parint,
                              int64_t S, int64_t E, double* out);

/* ModelMap::scale_map_by_solid_angle / _by_exposure /
 * _by_correction_factors (model_map.scale_map_by_*.cxx): in place on a
 * double [S, H, W, E] map. */
int xts_scale_map(xts_ctx* ctx, double* map, int64_t S, int64_t H, int64_t W,
                  int64_t E, const double* factor, int32_t kind);

/* SkyGeom::pix2sph / sph2pix (sky_geom/sky_geom.hxx:179-250) for n points,
 * evaluated on the device with the same WCS code the kernels use. */
int xts_pix2sph(xts_ctx* ctx, const xts_wcs* wcs, const double* pix /*[n,2]*/,
                int64_t n, double* sph /* [n, 2] */);
int xts_sph2pix(xts_ctx* ctx, const xts_wcs* wcs, const double* sph, int64_t n,
                double* pix);

/* ------------------------------------------------------ diffuse sources */
/* The diffuse-source model map: ModelMap::diffuse_src_model_map_wcs
 * (model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97; the reference
 * declares it but does not compile it, model_map/model_map.hxx:30 and
 * model_map/CMakeLists.txt:8-15).  A sky cube (the galactic diffuse
 * MapCube, fits::read_allsky_cropped's input) is cropped to the box of the
 * model map (fits/fits.read_allsky_cropped.cpp:150-216), resampled by
 * corner-aligned trilinear interpolation onto a grid of target_pix_size
 * pixels (sky_geom/sky_geom.hxx:350-380, math/trilerp.hpp:9-56) and across
 * energy planes, convolved per energy with the PSF kernel of the diffuse
 * source (psf/psf.make_kernel.cpp:8-90) through FFTs
 * (model_map/mm_fft/fft.convolve_map_psf.cpp:8-39), sampled bilinearly at
 * the model map's pixel centres and scaled by pixel solid angle and
 * exposure (diffuse_src_model_map_wcs.cxx:76, 88). */

/* The crop, grid, kernel and transform sizes of a diffuse map. */
typedef struct {
    int64_t crop_a1, crop_a2;   /* first cube pixel of the crop (1-based; a1
                                   may lie outside 1..naxis1 when wrapped) */
    int64_t crop_n1, crop_n2;   /* crop size in cube pixels */
    int64_t m1, m2;             /* resampled grid size */
    double step1, step2;        /* cube pixels per resampled pixel */
    xts_wcs grid;               /* WCS of the resampled grid */
    int64_t wrap;               /* 1: all-sky CAR cube, longitude wraps */
    int64_t nk;                 /* kernel size (odd, min(m1, m2)) */
    double kc1, kc2;            /* kernel centre on the grid (1-based pixels):
                                   the model map's reference direction */
    int64_t fft_p, fft_q;       /* transform lengths along axis 1 and 2 */
} xts_diffuse_geom;

/* The MapCube the diffuse map reads (fits::read_allsky_cropped's input,
 * fits/fits.read_allsky_cropped.cpp:13-41, 150-216): the primary image of
 * a FITS file, BITPIX -32 or -64, NAXIS 2 or 3, CAR or AIT.  wcs gets
 * readWcsParams's keywords (CROTA2 0 when absent; is_galactic when CTYPE1
 * starts with GLON); planes NAXIS3 (1 for a 2-D image); data (optional)
 * [planes][naxis2][naxis1] float; energies (optional) the first column of
 * the ENERGIES table, *n_energies entries (0 without the table).  Call
 * with data and energies NULL for the sizes. */
int xts_read_mapcube(const char* path, xts_wcs* wcs, int64_t* planes, float* data,
                     double* energies, int64_t* n_energies);

/* Host only: the geometry xts_diffuse_src_model_map_wcs_dev uses.
 * margin_deg widens the crop on every side (0: the map's own box plus one
 * pixel). */
int xts_diffuse_geometry(const xts_wcs* roi, const xts_wcs* cube, double target_pix_size,
                         double margin_deg, xts_diffuse_geom* out);

/* Device pointers.  cube: FP32 [n_planes, cube.naxis2, cube.naxis1] (FITS
 * order, axis 1 fastest); upsf: the diffuse source's PSF table [401, E]
 * (uPsf[0] of the reference's [S, 401, E]); expo: [E]; out: [E, H, W] with
 * H = roi.naxis1, W = roi.naxis2 (the model maps' layout), double, or float
 * with XTS_OUT_F32.  A model map in the celestial frame over a galactic cube
 * (or the reverse) converts coordinates through the J2000 galactic frame;
 * pixels whose centres fall outside the cube are 0. */
int xts_diffuse_src_model_map_wcs_dev(xts_ctx* ctx, const xts_wcs* roi, const xts_wcs* cube_wcs,
                                      const float* cube, int64_t n_planes, const double* upsf,
                                      const double* expo, int64_t E, double target_pix_size,
                                      double margin_deg, void* out, int32_t flags);
/* The same with host pointers; out double. */
int xts_diffuse_src_model_map_wcs(xts_ctx* ctx, const xts_wcs* roi, const xts_wcs* cube_wcs,
                                  const float* cube, int64_t n_planes, const double* upsf,
                                  const double* expo, int64_t E, double target_pix_size,
                                  double margin_deg, double* out);

/* Psf::make_kernel (psf/psf.make_kernel.cpp:8-90), host pointers: out
 * [E, nk, nk], nk odd; kernel pixel (h, w) is the grid's pixel (kc1 + h - c,
 * kc2 + w - c), c = (nk - 1) / 2, valued by the table at its separation from
 * the grid direction at (kc1, kc2) and normalised to sum 1 per energy.  The
 * reference's call is the case of a grid whose CRPIX is (c + 1, c + 1) and
 * (kc1, kc2) = CRPIX. */
int xts_psf_make_kernel(xts_ctx* ctx, int64_t nk, const xts_wcs* grid, double kc1, double kc2,
                        const double* upsf, int64_t E, double* out);

/* FFT::convolve_map_psf (model_map/mm_fft/fft.convolve_map_psf.cpp:8-39),
 * host pointers: image [C, H, W], kernel [C, KH, KW] -> out [C, H + KH - 1,
 * W + KW - 1], the full linear convolution of each plane with its kernel. */
int xts_fft_convolve_map_psf(xts_ctx* ctx, const double* image, int64_t C, int64_t H, int64_t W,
                             const double* kernel, int64_t KH, int64_t KW, double* out);

/* ------------------------------------------------------ learned surrogate */
/* The surrogate of the far-pixel cubature: a 4 -> 64 -> 64 -> 64 -> 1 SiLU
 * network that predicts, from four features of a pixel's point separations
 * about the one table node they straddle, the excess B1 - |A| of their
 * absolute first moment over the first moment's absolute value (the part of
 * the pixel's row weights that is not smooth in its geometry).  No
 * counterpart in the reference, which runs the point loop of
 * mm_cubature/convolve_pixel_psf_logsep.hpp:23-96 for every pixel.
 * params: FP32 [XTS_MLP_PARAMS] = W1 [64][4], b1 [64], W2 [64][64], b2 [64],
 * W3 [64][64], b3 [64], w4 [64], b4; norm: FP32 [XTS_MLP_NORM] = mu [4],
 * 1/sd [4], y scale, y offset (x' = (x - mu)/sd, y = y' scale + offset).
 * All pointers are device pointers; calls enqueue on the context stream. */
#define XTS_MLP_PARAMS 8705
#define XTS_MLP_NORM 10

/* Training samples from the FP64 cubature: n_candidates random (source,
 * pixel) pairs of the image (sources over the image widened by a quarter on
 * each side); those the model map's row-weight pass would give the network
 * are written compactly as rows of 8 floats (sigma_c, g_u, g_v, kappa,
 * target B1 - |A|, A, centre separation in degrees, weight total), at most
 * capacity rows; *count (device uint64, cleared by the caller) receives the
 * number found.  rule [N, 3] device doubles, as for the model map. */
int xts_surrogate_samples_dev(xts_ctx* ctx, const xts_wcs* wcs, const double* rule, int64_t N,
                              uint64_t seed, int64_t n_candidates, int64_t capacity, float* out,
                              uint64_t* count);
/* FP32 forward of B samples (x rows of ldx floats, features first) -> y [B]. */
int xts_mlp_forward_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                        int64_t ldx, int64_t B, float* y);
/* BF16 inference on the matrix cores (high and low BF16 weight parts,
 * FP32 accumulation) -> y [B]. */
int xts_mlp_infer_bf16_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                           int64_t ldx, int64_t B, float* y);
/* FP8 inference (the record of XTS_SURROGATE_FP8: e4m3 hidden layers with
 * per-layer weight scales and per-wave activation scales, FP32
 * accumulation) -> y [B]. */
int xts_mlp_infer_fp8_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                          int64_t ldx, int64_t B, float* y);
/* Gradient of the mean squared error of the normalised outputs over B
 * samples (targets y_target, stride ldy) -> grad [XTS_MLP_PARAMS], loss [1]. */
int xts_mlp_grad_dev(xts_ctx* ctx, const float* params, const float* norm, const float* x,
                     int64_t ldx, const float* y_target, int64_t ldy, int64_t B, float* grad,
                     float* loss);
/* One Adam step (step >= 1) on params [XTS_MLP_PARAMS]; grad is cleared. */
int xts_mlp_adam_dev(xts_ctx* ctx, float* params, float* grad, float* m, float* v, int64_t step,
                     float lr, float beta1, float beta2, float eps);
/* steps Adam steps of minibatches of `batch` rows drawn from the n samples,
 * learning rate lr with cosine decay to lr/20; loss_hist [steps] (or NULL)
 * receives each step's mean squared error. */
int xts_surrogate_train_dev(xts_ctx* ctx, float* params, const float* norm, const float* x,
                            int64_t ldx, const float* y, int64_t ldy, int64_t n, int64_t steps,
                            int64_t batch, float lr, uint64_t seed, float* loss_hist);
/* The network used by model maps with XTS_SURROGATE (NULL params clears it);
 * packed into the context on its stream and synchronised. */
int xts_ctx_set_surrogate(xts_ctx* ctx, const float* params, const float* norm);

/* ------------------------------------------------------------- output */
/* Fermi::fits::write_src_model (fits/fits.write_src_model.cxx:10-66): a FITS
 * file with a PRIMARY HDU without data (HDUNAME = 'PRIMARY') and one IMAGE
 * HDU per source, EXTNAME = names[s], holding the source's cube in FITS axis
 * order NAXIS1 = H (the image's first, longitude axis), NAXIS2 = W,
 * NAXIS3 = E, with CHECKSUM and DATASUM cards.  map: double [S, H, W, E]
 * (XTS_LAYOUT_SHWE, the reference's) or [S, E, H, W] (XTS_LAYOUT_SEHW);
 * XTS_OUT_F32 writes BITPIX -32 (gtsrcmaps' type), otherwise -64 (the
 * reference's).  wcs, if not NULL, adds the image's celestial WCS cards to
 * each image HDU.  The reference declares NAXIS1 = W and NAXIS2 = H for the
 * same data (the same file for square images); this writer keeps NAXIS1 as
 * the fastest data axis.  Host only: needs no context or device. */
int xts_write_src_model(const char* path, const double* map, int64_t S, int64_t H,
                        int64_t W, int64_t E, const char* const* names, int32_t flags,
                        const xts_wcs* wcs);

/* ------------------------------------------------------------- input */
/* fits::read_irf_pars (fits/fits.read_irf_pars.cxx:12-126): the binary
 * table named tblname (EXTNAME, case-insensitive) whose columns are all
 * float32 ('E'), as rows of the columns' elements concatenated in column
 * order.  With rowdata NULL it only reports the number of columns, their
 * repeat counts (the first cap_cols go to extents), the number of rows and
 * the row width; then call again with rowdata of nrows x row_width floats.
 * A table with a column of another type is an error, as in the reference.
 * Host only. */
int xts_read_irf_pars(const char* path, const char* tblname, int64_t* ncols,
                      int64_t* extents, int64_t cap_cols, int64_t* nrows,
                      int64_t* row_width, float* rowdata);

/* Livetime-cube table keywords: Obs::ExposureCubeData
 * (observation/obs_types.hxx:10-20) without the arrays. */
typedef struct {
    int64_t nside, nbrbins;
    double cosmin;
    int32_t thetabin; /* THETABIN other than 'COSTHETA' */
    char ordering[16], coordsys[16];
} xts_expcube_info;

/* fits::read_expcube (fits/fits.read_expcube.cxx:8-116): the keywords of the
 * HEALPix livetime table tblname ("EXPOSURE", "WEIGHTED_EXPOSURE") and, for
 * each non-NULL pointer, its columns 1-3 read as float: cosbins
 * [12 nside^2 * nbrbins] (pixel-major), ra and dec [12 nside^2].  Only
 * ORDERING = NESTED and COORDSYS = EQU cubes are accepted, as in the
 * reference.  Host only. */
int xts_read_expcube(const char* path, const char* tblname, xts_expcube_info* info,
                     float* cosbins, float* ra, float* dec);

#ifdef __cplusplus
}
#endif
#endif
