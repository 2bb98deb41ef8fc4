// IRF math on the path (reference: irf/irf.hxx:14-84,
// irf/irf.normalize_rpsf.cpp:11-61, irf/irf.prepare_grid.cpp:26-102).
#pragma once

#include <vector>

#include "xtsrcmaps/irf/irf_types.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::Irf {

auto livetime_efficiency_factors(std::vector<double> const& logEs, IrfEffic const& effic)
    -> Tensor<double, 2>;

auto aeff_value(std::vector<double> const& costhet, std::vector<double> const& logEs,
                IrfData3 const& AeffData) -> Tensor<double, 2>;

// King normalisation and SCORE/STAIL scaling of one IRF section, in place.
void normalize_rpsf(psf::Data& psfdata);

// One FITS IRF row (energy lo/hi, cos theta lo/hi, Ng parameter blocks of
// [Mc0, Me0]) -> the padded grid.
auto prepare_grid(std::vector<float> const& e_lo, std::vector<float> const& e_hi,
                  std::vector<float> const& c_lo, std::vector<float> const& c_hi,
                  std::vector<float> const& params, std::size_t n_grids) -> IrfData3;

}  // namespace Fermi::Irf
