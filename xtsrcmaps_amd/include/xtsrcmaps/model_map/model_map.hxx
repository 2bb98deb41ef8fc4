// (c) Point-source model map (reference: model_map/model_map.hxx:11-74).
// Maps are [S, H, W, E] FP64 host tensors as in the reference.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/psf/psf.hxx"
#include "xtsrcmaps/sky_geom/sky_geom.hxx"
#include "xtsrcmaps/source/source.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::ModelMap {

auto point_src_model_map_wcs(size_t Nh, size_t Nw, SkyGeom<double> const& skygeom,
                             Tensor<double, 2> const& src_sph,
                             std::vector<std::string> const& src_names,
                             Tensor<double, 2> const& exposures, Tensor<double, 3> const& uPsf,
                             Tensor<double, 3> const& partial_integrals) -> Tensor<double, 4>;

void scale_map_by_solid_angle(Tensor<double, 4>& model_map, SkyGeom<double> const& skygeom);
void scale_map_by_exposure(Tensor<double, 4>& model_map, Tensor<double, 2> const& exposures);
auto map_correction_factor(Tensor<double, 2> const& inv_map_integ,
                           Tensor<double, 1> const& psf_radius,
                           std::vector<bool> const& is_in_fov, Tensor<double, 3> const& mean_psf,
                           Tensor<double, 3> const& partial_integrals) -> Tensor<double, 2>;
void scale_map_by_correction_factors(Tensor<double, 4>& model_map,
                                     Tensor<double, 2> const& factor);
auto psf_boundary_radius(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sph,
                         SkyGeom<double> const& skygeom)
    -> std::pair<Tensor<double, 1>, std::vector<bool>>;
auto map_integral(Tensor<double, 4> const& model_map, Tensor<double, 2> const& src_sph,
                  SkyGeom<double> const& skygeom, Tensor<double, 1> const& psf_radius,
                  std::vector<bool> const& is_in_fov) -> Tensor<double, 2>;

// ModelMap::diffuse_src_model_map_wcs (model_map.diffuse_src_model_map_wcs.cxx:8-97,
// declared but not compiled by the reference): src_names[0] is the MapCube
// FITS file (the galactic diffuse cube), exposures [S, E] and uPsf [S, 401,
// E] the diffuse source's (row 0: at the image's reference direction); the
// cube is cropped, resampled to target_pix_size, convolved per energy with
// the PSF kernel by FFTs, sampled at the image's pixel centres and scaled by
// solid angle and exposure.  Returns [1, Nh, Nw, E].
auto diffuse_src_model_map_wcs(size_t const Nh, size_t const Nw, SkyGeom<double> const& skygeom,
                               Tensor<double, 2> const& src_sph,
                               std::vector<std::string> const& src_names,
                               Tensor<double, 2> const& exposures, Tensor<double, 3> const& uPsf,
                               Tensor<double, 3> const& partial_integrals,
                               double const target_pix_size) -> Tensor<double, 4>;

// model_map.compute_src_maps.cxx:6-22
template <Source::SourceConcept T>
auto compute_srcmaps(Obs::XtObs const& obs, Source::SourceData<T> src,
                     Exposure::XtExp const& exp, Psf::XtPsf const& psf) -> Tensor<double, 4> {
    return point_src_model_map_wcs((size_t)obs.Nh, (size_t)obs.Nw, obs.skygeom, src.sph_locs,
                                   src.names, exp.exposure, psf.uPsf, psf.partial_psf_integral);
}

}  // namespace Fermi::ModelMap

namespace Fermi::FFT {
// FFT::convolve_map_psf (model_map/mm_fft/fft.convolve_map_psf.cpp:8-39): the
// full linear convolution of each plane of image [C, H, W] with the same
// plane of kernel [C, KH, KW]: [C, H + KH - 1, W + KW - 1].
auto convolve_map_psf(Tensor<double, 3> const& image, Tensor<double, 3> const& kernel)
    -> Tensor<double, 3>;
}  // namespace Fermi::FFT
