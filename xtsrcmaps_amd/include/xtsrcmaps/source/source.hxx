// Source data crossing into the model-map path (reference:
// source/source.hxx:93-131): only the point-source positions and names.
#pragma once

#include <string>
#include <vector>

#include "xtsrcmaps/tensor/tensor.hpp"

namespace Fermi::Source {

struct PointSource {
    std::string name = "";
    float roi_center_dist = 0.0;
};

template <typename T>
concept SourceConcept = std::same_as<T, PointSource>;

template <SourceConcept T>
struct SourceData {
    std::vector<T> srcs;
    Fermi::Tensor<double, 2> sph_locs;  // [S, 2] (RA, DEC) degrees
    std::vector<std::string> names;
};

}  // namespace Fermi::Source
