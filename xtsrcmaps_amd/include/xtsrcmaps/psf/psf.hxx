// (b) Mean PSF (reference: psf/psf.hxx:12-103).
#pragma once

#include <array>
#include <cmath>
#include <tuple>
#include <vector>

#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/irf/irf_types.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/sky_geom/sky_geom.hxx"
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

// separation of fractional node index x and its inverse (psf/psf.hxx:33-43)
inline auto separation(double x) -> double {
    const double delta = std::exp(std::log(7e5) / 399.);
    return 1e-4 * (x >= 1. ? std::pow(delta, x - 1.) : x);
}
inline auto inverse_separation(double s) -> double {
    return s < 1e-4 ? 1e4 * s : 1. + std::log(s * 1e4) / (std::log(7e5) / 399.);
}

auto fast_separation_lower_index(Tensor<double, 1> seps) -> Tensor<int, 1>;
auto bilerp(std::vector<double> const& costhetas, std::vector<double> const& logEs,
            Tensor<double, 1> const& par_cosths, Tensor<double, 1> const& par_logEs,
            Tensor<double, 3> const& kings) -> Tensor<double, 3>;  // [C, 401, E]
auto corrected_exposure_psf(Tensor<double, 3> const& obs_psf,  // [C, 401, E]
                            Tensor<double, 2> const& obs_aeff,  // [C, E]
                            Tensor<double, 2> const& src_exposure_cosbins,
                            Tensor<double, 2> const& src_weighted_exposure_cosbins,
                            Tensor<double, 2> const& front_LTF) -> Tensor<double, 3>;
auto mean_psf(Tensor<double, 3> const& front_corrected_psf,
              Tensor<double, 3> const& back_corrected_psf,
              Tensor<double, 2> const& exposures) -> Tensor<double, 3>;  // [S, 401, E]
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

// Psf::make_kernel (psf/psf.make_kernel.cpp:8-90): the PSF of uPsf's first
// source [S, 401, E] on an Nk x Nk block (Nk odd) of skygeom's pixel grid
// centred at its reference pixel (CRPIX), valued at each pixel's separation
// from the reference direction and normalised to sum 1 per energy: [E, Nk,
// Nk].  The reference's block starts at pixel (1, 1), which is the same
// block when CRPIX is ((Nk + 1) / 2, (Nk + 1) / 2); its two table nodes are
// blended as the point-source pixel integral blends them.
auto make_kernel(size_t const Nk, SkyGeom<double> const& skygeom,
                 Tensor<double, 3> const& uPsf) -> Tensor<double, 3>;

}  // namespace Fermi::Psf
