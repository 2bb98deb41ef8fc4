// main_calls.cpp -- the reference's call lines into the model-map path
// (xtsrcmaps/main.cxx:14-31), unchanged, against this library's headers.
// Compiled with -fsyntax-only by tests/test_boundary_cpu.py; never linked.
//
// The modules outside the path keep their own headers in the reference
// (cli/cli.hxx:4, observation/observation.hxx:8); their two declarations
// are repeated here because those headers are not part of this library.
#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/irf/irf.hxx"
#include "xtsrcmaps/model_map/model_map.hxx"
#include "xtsrcmaps/psf/psf.hxx"
#include "xtsrcmaps/source/source.hxx"

namespace Fermi::Config {
auto parse_cli_to_cfg(int const argc, char** argv) -> Config::XtCfg;
}
namespace Fermi::Obs {
auto collect_observation_data(Fermi::Config::XtCfg const&) -> Obs::XtObs;
}

using namespace Fermi;

int main(int const argc, char** argv) {
    struct ModelMaps {
        Tensor<double, 4> pt = {};
        Tensor<double, 4> ds = {};
    };

    auto const cfg = Config::parse_cli_to_cfg(argc, argv);
    auto const obs = Obs::collect_observation_data(cfg);
    auto const src = Source::collect_source_model(cfg, obs);
    auto const irf = Irf::collect_irf_data(cfg, obs);
    auto maps = ModelMaps{};

    if (src.point.srcs.size()) {
        auto ptexp = Exposure::compute_exposure(cfg, obs, src.point, irf);
        auto ptpsf = Psf::compute_psf_data(obs, irf, ptexp);
        maps.pt = ModelMap::compute_srcmaps(obs, src.point, ptexp, ptpsf);
    }
    return 0;
}
