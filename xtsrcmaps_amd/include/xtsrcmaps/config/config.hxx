// Run configuration passed into the path (reference: config/config.hxx:9-46).
// The model-map path reads none of its fields; the struct is carried so that
// Exposure::compute_exposure keeps the reference's four-argument signature.
#pragma once

#include <optional>
#include <string>

namespace Fermi::Config {

struct XtCfg {
    std::string sctable = "SC_DATA";
    std::string expcube = "analysis/3C279_binned_ltcube.fits";
    std::string cmap = "analysis/3C279_binned_ccube.fits";
    std::string srcmdl = "analysis/3C279_input_model_point.xml";
    std::string bexpmap = "analysis/3C279_binned_allsky_expcube.fits";
    std::string wmap = "none";
    std::string outfile = "analysis/3C279_binned_srcmaps_local.fits";
    std::string psf_file = "analysis/psf_P8R3_SOURCE_V2_FB.fits";
    std::string aeff_file = "analysis/aeff_P8R3_SOURCE_V2_FB.fits";
};

// The parameter-file reader and validation live in the reference's config
// module (config/config.cxx), which is outside the path and not replaced.
auto parse_parfile(const std::string& filename) -> std::optional<XtCfg>;
auto validate_cfg(XtCfg const& ocfg) -> XtCfg;

}  // namespace Fermi::Config
