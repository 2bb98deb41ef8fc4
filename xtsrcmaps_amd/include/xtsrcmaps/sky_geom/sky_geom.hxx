// Fermi::SkyGeom -- WCS geometry of a model-map image (reference:
// sky_geom/sky_geom.hxx:17-349).  Same constructors and the methods the
// model-map path calls; the projection math (CAR and AIT, Calabretta &
// Greisen 2002, CROTA2 convention of Greisen & Calabretta 2002) is this
// library's own, shared with its GPU kernels, in place of wcslib.
#pragma once

#include <array>
#include <cmath>
#include <string>
#include <utility>

#include "xtsrcmaps_amd.h"

namespace Fermi {

struct WcsConfig {
    std::array<long, 3> naxes;    // NAXIS extent of image
    std::array<double, 3> crpix;  // reference pixel
    std::array<double, 3> crval;  // world coordinate at the reference pixel
    std::array<double, 3> cdelt;
    double axis_rot;
    std::string proj_name;        // "CAR" or "AIT"
    bool is_galactic;
};

template <typename T>
class SkyGeom {
  public:
    SkyGeom() = default;
    SkyGeom(long naxis1, long naxis2, std::string const& proj_name, bool is_galactic,
            double crpix1, double crpix2, double crval1, double crval2, double cdelt1,
            double cdelt2, double axis_rot);
    explicit SkyGeom(WcsConfig const& cfg)
        : SkyGeom(cfg.naxes[0], cfg.naxes[1], cfg.proj_name, cfg.is_galactic, cfg.crpix[0],
                  cfg.crpix[1], cfg.crval[0], cfg.crval[1], cfg.cdelt[0], cfg.cdelt[1],
                  cfg.axis_rot) {}

    long const* naxes() const { return naxes_.data(); }
    double const* crval() const { return crval_.data(); }
    double const* crpix() const { return crpix_.data(); }
    double const* cdelt() const { return cdelt_.data(); }
    double axis_rot() const { return wcs_.crota2; }

    auto sph2pix(std::array<T, 2> const& ss) const -> std::array<T, 2>;
    auto sph2pix(std::pair<T, T> const& ss) const -> std::pair<T, T>;
    auto pix2sph(std::array<T, 2> const& px) const -> std::array<T, 2>;
    auto pix2sph(T first, T second) const -> std::array<T, 2>;
    auto dir2sph(std::array<T, 3> const& dir) const -> std::array<T, 2>;
    auto pix2dir(std::array<T, 2> const& px) const -> std::array<T, 3>;
    auto sph2dir(std::pair<T, T> const& s) const -> std::array<T, 3>;
    auto sph2dir(std::array<T, 2> const& s) const -> std::array<T, 3> {
        return sph2dir(std::pair<T, T>{s[0], s[1]});
    }
    auto refpix() const -> std::array<T, 2> { return {crpix_[0], crpix_[1]}; }
    auto refsph() const -> std::array<T, 2> { return {crval_[0], crval_[1]}; }
    auto refdir() const -> std::array<T, 3> { return sph2dir(refsph()); }
    auto srcpixoff(std::array<T, 3> const& src, std::array<T, 3> const& pix) const -> T {
        return dir_diff(src, pix) * T(57.2957795130823208767981548141051703324);
    }
    static auto dir_diff(std::array<T, 3> const& L, std::array<T, 3> const& R) -> T {
        const T x = L[0] - R[0], y = L[1] - R[1], z = L[2] - R[2];
        return T(2) * std::asin(T(0.5) * std::sqrt(x * x + y * y + z * z));
    }
    static auto pix_diff(std::array<T, 2> const& L, std::array<T, 2> const& R,
                         SkyGeom const& skygeom) -> T {
        return dir_diff(skygeom.pix2dir(L), skygeom.pix2dir(R));
    }
    auto wcs_config() const -> WcsConfig;

    // the C-ABI description of this geometry
    xts_wcs const& c_wcs() const { return wcs_; }

  private:
    xts_wcs wcs_{};
    std::array<long, 2> naxes_{};
    std::array<double, 3> crval_{}, crpix_{}, cdelt_{};
    std::string proj_name_;
};

extern template class SkyGeom<double>;

}  // namespace Fermi
