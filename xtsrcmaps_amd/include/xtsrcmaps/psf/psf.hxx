// (b) Mean PSF (reference: psf/psf.hxx:12-103).
#pragma once

#include <array>
#include <cmath>
#include <tuple>
#include <vector>

#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/irf/irf_types.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::Psf {

struct XtPsf {
    Tensor<double, 3> uPsf;                  // [S, 401, E] normalised
    Tensor<double, 3> partial_psf_integral;  // [S, 401, E]
};

constexpr unsigned short sep_arr_len = 401;
using SepArr = std::array<double, sep_arr_len>;

// 0, then 1e-4 deg .. 70 deg in 400 log steps (psf/psf.hxx:24-31)
inline auto separations(double xmin = 1e-4, double xmax = 70.) -> SepArr {
    SepArr sep{};
    const double xstep = std::log(xmax / xmin) / 399.;
    for (std::size_t i = 0; i < 400; ++i) sep[i + 1] = xmin * std::exp(i * xstep);
    return sep;
}

auto fast_separation_lower_index(Tensor<double, 1> seps) -> Tensor<int, 1>;
auto king(Irf::psf::Data const& data) -> Tensor<double, 3>;  // [401, Mc, Me]
auto partial_total_integral(Tensor<double, 3> const& mean_psf)
    -> std::pair<Tensor<double, 3>, Tensor<double, 2>>;
auto normalize(Tensor<double, 3>& mean_psf, Tensor<double, 2> const& total_integrals) -> void;

auto psf_lookup_table_and_partial_integrals(
    Irf::psf::Pass8FB const& data, std::vector<double> const& costhetas,
    std::vector<double> const& logEs, Tensor<double, 2> const& front_aeff,
    Tensor<double, 2> const& back_aeff, Tensor<double, 2> const& src_exposure_cosbins,
    Tensor<double, 2> const& src_weighted_exposure_cosbins, Tensor<double, 2> const& front_LTF,
    Tensor<double, 2> const& exposure) -> std::tuple<Tensor<double, 3>, Tensor<double, 3>>;

auto compute_psf_data(Obs::XtObs const& obs, Irf::XtIrf const& irf,
                      Exposure::XtExp const& exp) -> XtPsf;

}  // namespace Fermi::Psf
