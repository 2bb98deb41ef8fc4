// (a) Exposure (reference: exposure/exposure.hxx:11-88).
#pragma once

#include <vector>

#include "xtsrcmaps/config/config.hxx"
#include "xtsrcmaps/irf/irf.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/source/source.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::Exposure {

struct ExposureMap {
    std::size_t nside;
    std::size_t nbins;
    Tensor<double, 2> params;
};

struct XtExp {
    std::vector<double> exp_costhetas;
    Tensor<double, 2> front_aeff;
    Tensor<double, 2> back_aeff;
    Tensor<double, 2> src_exposure_cosbins;
    Tensor<double, 2> src_weighted_exposure_cosbins;
    Tensor<double, 2> exposure;
};

auto map(Obs::ExposureCubeData const&) -> ExposureMap;
auto costhetas(Obs::ExposureCubeData const&) -> std::vector<double>;
auto contract(Tensor<double, 2> const& A, Tensor<double, 2> const& B) -> Tensor<double, 2>;
auto src_cosbins(Tensor<double, 2> const&, ExposureMap const&) -> Tensor<double, 2>;
auto exposure(Tensor<double, 2> const& src_exposure_cosbins,
              Tensor<double, 2> const& src_weighted_exposure_cosbins,
              Tensor<double, 2> const& front_aeff, Tensor<double, 2> const& back_aeff,
              Tensor<double, 2> const& front_ltfs) -> Tensor<double, 2>;

// The configuration is not read (as in the reference); it is kept so that
// main.cxx:28's call compiles unchanged.
template <Source::SourceConcept T>
auto compute_exposure(Config::XtCfg const& /*cfg*/, Obs::XtObs const& obs,
                      Source::SourceData<T> const& src, Irf::XtIrf const& irf) -> XtExp {
    auto const exp_costhetas = costhetas(obs.exp_cube);
    auto const exp_map = map(obs.exp_cube);
    auto const wexp_map = map(obs.weighted_exp_cube);
    auto const ecb = src_cosbins(src.sph_locs, exp_map);
    auto const wecb = src_cosbins(src.sph_locs, wexp_map);
    auto const front_aeff =
        Irf::aeff_value(exp_costhetas, obs.logEs, irf.aeff_irf.front.effective_area);
    auto const back_aeff =
        Irf::aeff_value(exp_costhetas, obs.logEs, irf.aeff_irf.back.effective_area);
    auto const expo = exposure(ecb, wecb, front_aeff, back_aeff, irf.front_LTF);
    return {exp_costhetas, front_aeff, back_aeff, ecb, wecb, expo};
}

}  // namespace Fermi::Exposure
