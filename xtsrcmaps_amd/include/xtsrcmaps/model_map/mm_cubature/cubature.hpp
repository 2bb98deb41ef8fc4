// Pixel cubature (reference: model_map/mm_cubature/cubature.hpp:12-17).
#pragma once

#include <vector>

#include "xtsrcmaps/sky_geom/sky_geom.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi {

// Raw cubature of the PSF table over every pixel -> [S, H, W, E], with the
// reference's production rule (Festa-Sommariva degree 7, 12 points).
auto cubature_convolve_map_psf(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sphcrds,
                               Tensor<double, 3> const& psf_lut,
                               SkyGeom<double> const& skygeom) -> Tensor<double, 4>;

// The same with an explicit rule: rows (u, v, w) on [-1, 1]^2, sum w = 4.
auto cubature_convolve_map_psf(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sphcrds,
                               Tensor<double, 3> const& psf_lut, SkyGeom<double> const& skygeom,
                               std::vector<double> const& rule_uvw) -> Tensor<double, 4>;

}  // namespace Fermi
