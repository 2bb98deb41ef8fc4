// FITS input and output of the path (reference: fits/fits.hxx:17-64):
// the IRF tables, the livetime cube and the source-map file.
#pragma once

#include <optional>
#include <string>
#include <vector>

#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::fits {

// One IRF table (fits/fits.hxx:44-48): column repeat counts, their offsets
// in a row, and the rows as float.
struct TablePars {
    std::vector<size_t> extents;
    std::vector<size_t> offsets;
    Tensor<float, 2> rowdata;
};

// fits/fits.read_irf_pars.cxx:12-126.  std::nullopt (with a message on
// stderr) when the file or table cannot be read or a column is not float32.
auto read_irf_pars(std::string const& filename, std::string tblname)
    -> std::optional<TablePars>;

// fits/fits.read_expcube.cxx:8-116.  std::nullopt (with a message on stderr)
// on failure, or unless ORDERING = NESTED and COORDSYS = EQU.
auto read_expcube(std::string const& filename, std::string const& tblname)
    -> std::optional<Obs::ExposureCubeData>;

// One IMAGE HDU per source (EXTNAME = its name) of the [S, H, W, E] map,
// FP64, with CHECKSUM/DATASUM, after a PRIMARY HDU without data.  The cube
// is stored in FITS axis order NAXIS1 = H, NAXIS2 = W, NAXIS3 = E.  Throws
// std::runtime_error when the file cannot be written, std::invalid_argument
// when the names do not match the sources.
auto write_src_model(std::string const& filename, Tensor<double, 4>& model_map,
                     std::vector<std::string> const& srcs) -> void;

}  // namespace Fermi::fits
