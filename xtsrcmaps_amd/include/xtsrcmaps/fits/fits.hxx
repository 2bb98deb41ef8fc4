// Source-map output (reference: fits/fits.hxx:56-58, fits.write_src_model.cxx).
#pragma once

#include <string>
#include <vector>

#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::fits {

// One IMAGE HDU per source (EXTNAME = its name) of the [S, H, W, E] map,
// FP64, with CHECKSUM/DATASUM, after a PRIMARY HDU without data.  The cube
// is stored in FITS axis order NAXIS1 = H, NAXIS2 = W, NAXIS3 = E.  Throws
// std::runtime_error when the file cannot be written, std::invalid_argument
// when the names do not match the sources.
auto write_src_model(std::string const& filename, Tensor<double, 4>& model_map,
                     std::vector<std::string> const& srcs) -> void;

}  // namespace Fermi::fits
