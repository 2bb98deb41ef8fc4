// Source data crossing into the model-map path (reference:
// source/source.hxx:14-131).  The structs keep the reference's fields so
// that its XML source-model parser (source/source.parse_src_mdl.cxx, outside
// the path) fills them unchanged; the path itself reads only the positions
// (sph_locs) and names of a SourceData.
#pragma once

#include <algorithm>
#include <array>
#include <concepts>
#include <iterator>
#include <string>
#include <variant>
#include <vector>

#include "xtsrcmaps/config/config.hxx"
#include "xtsrcmaps/observation/obs_types.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::Source {

struct SourceParameter {
    std::string name = "";
    unsigned long free = 0ul;
    float value = 0.0f;
    float scale = 1.0f;
    float min = 0.0f;
    float max = 0.0f;
};

enum class SpectrumType {
    BPLExpCutoff,
    BrokenPowerLaw,
    BrokenPowerLaw2,
    ConstantValue,
    ExpCutoff,
    FileFunction,
    Gaussian,
    LogParabola,
    PowerLaw,
    PowerLaw2,
    PLSuperExpCutoff2,
    PLSuperExpCutoff3,
    PLSuperExpCutoff4,
    Unknown,
};

struct NormalSpectrum {
    std::vector<SourceParameter> params = {};
    SpectrumType type = SpectrumType::Unknown;
};

struct FileSpectrum {
    std::string file = "";
    std::vector<SourceParameter> params = {};
    bool edisp = false;
};

using Spectrum = std::variant<NormalSpectrum, FileSpectrum>;

struct ConstantValueSpatialModel {
    std::vector<SourceParameter> params = {};
};

struct MapCubeFunctionSpatialModel {
    std::string file = "";
    std::vector<SourceParameter> params = {};
};

struct SkyDirFunctionSpatialModel {
    std::vector<SourceParameter> params = {};
};

struct UnknownSpatialModel {};

using SpatialModel = std::variant<ConstantValueSpatialModel, MapCubeFunctionSpatialModel,
                                  SkyDirFunctionSpatialModel, UnknownSpatialModel>;

struct PointSource {
    std::string name = "";
    Spectrum spectrum = {};
    SpatialModel spatial_model = {};
    float roi_center_dist = 0.0;
};

struct DiffuseSource {
    std::string name = "";
    Spectrum spectrum = {};
    SpatialModel spatial_model = {};
};

struct CompositeSource {
    std::string name = "";
    Spectrum spectrum = {};
    SpatialModel spatial_model = {};
};

struct UnknownSource {
    std::string name = "";
};

using Source = std::variant<PointSource, DiffuseSource, CompositeSource, UnknownSource>;

template <typename T>
concept SourceConcept = std::same_as<T, PointSource> || std::same_as<T, DiffuseSource> ||
                        std::same_as<T, CompositeSource> || std::same_as<T, UnknownSource>;

template <SourceConcept T>
struct SourceData {
    std::vector<T> srcs;
    Fermi::Tensor<double, 2> sph_locs;  // [S, 2] (RA, DEC) degrees
    std::vector<std::string> names;
};

struct XtSrc {
    SourceData<PointSource> point;
    SourceData<DiffuseSource> diffuse;
};

template <SourceConcept T>
auto source_names(std::vector<T> const& srcs) -> std::vector<std::string> {
    std::vector<std::string> names;
    names.reserve(srcs.size());
    std::transform(srcs.cbegin(), srcs.cend(), std::back_inserter(names),
                   [](auto const& s) -> std::string { return s.name; });
    return names;
}

struct SourceGroup {
    std::vector<PointSource> point;
    std::vector<DiffuseSource> diffuse;
};

// Declared as in the reference; defined by its source-model module
// (source/source.cxx, source/source.parse_src_mdl.cxx), which is outside the
// path and not replaced.
auto spherical_coords(std::vector<PointSource> const&) -> Fermi::Tensor<double, 2>;
auto refsph_coords(size_t const, std::array<double, 2> const) -> Fermi::Tensor<double, 2>;
auto parse_src_xml(std::string const& src_file_name) -> SourceGroup;
auto collect_source_model(Config::XtCfg const&, Obs::XtObs const&) -> XtSrc;

}  // namespace Fermi::Source
