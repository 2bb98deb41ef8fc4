// IRF math on the path (reference: irf/irf.hxx:14-84,
// irf/irf.normalize_rpsf.cpp:11-61, irf/irf.prepare_grid.cpp:26-102).
#pragma once

#include <optional>
#include <string>
#include <vector>

#include "xtsrcmaps/config/config.hxx"
#include "xtsrcmaps/fits/fits.hxx"
#include "xtsrcmaps/irf/irf_types.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
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

// The same from the table of read_irf_pars (irf.prepare_grid.cpp:26-102).
auto prepare_grid(fits::TablePars const& pars) -> IrfData3;

// irf/irf.cxx:24-68: the six tables of a P8R3 front/back IRF file, gridded;
// the PSF's King parameters normalised (irf.prepare_psf_data.cpp:6-25).
// std::nullopt when a table cannot be read.
auto load_aeff(std::string const& filename) -> std::optional<aeff::Pass8FB>;
auto load_psf(std::string const& filename) -> std::optional<psf::Pass8FB>;

// irf/irf.hxx:19: both IRF files of the configuration and the front LTF.
// Declared for main.cxx:23; its definition is the reference's own
// irf/irf.collect_irf_data.cpp, which calls only load_aeff, load_psf and
// livetime_efficiency_factors above and is kept as it is.
auto collect_irf_data(Config::XtCfg const& cfg, Obs::XtObs const& obs) -> XtIrf;

}  // namespace Fermi::Irf
