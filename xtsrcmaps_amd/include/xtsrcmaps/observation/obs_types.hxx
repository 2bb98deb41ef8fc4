// Observation data crossing into the model-map path (reference:
// observation/obs_types.hxx:11-32).
#pragma once

#include <string>
#include <vector>

#include "xtsrcmaps/sky_geom/sky_geom.hxx"

namespace Fermi::Obs {

struct ExposureCubeData {
    std::vector<float> cosbins;  // [12 nside^2, nbrbins] livetime, NESTED
    std::vector<float> ra;
    std::vector<float> dec;
    unsigned int nside = 0;
    unsigned int nbrbins = 40;
    double cosmin = 0.0;
    std::string ordering = "NESTED";
    std::string coordsys = "EQU";
    bool thetabin = false;
};

struct XtObs {
    long Nh;
    long Nw;
    std::vector<double> energies;
    std::vector<double> logEs;
    ExposureCubeData exp_cube;
    ExposureCubeData weighted_exp_cube;
    SkyGeom<double> skygeom;
};

}  // namespace Fermi::Obs
