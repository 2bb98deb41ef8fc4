// fermi_api.cpp -- the reference's C++ API of the model-map path
// (Fermi::Exposure, Fermi::Psf, Fermi::ModelMap, Fermi::Irf, SkyGeom) on
// host Tensors, implemented over the C ABI of include/xtsrcmaps_amd.h.
//
// Every compute function hands its arrays to the GPU entry point of the same
// name and returns a new Tensor (or scales one in place), as the reference
// does.  Failures raise std::runtime_error with the library's message; shape
// misuse raises std::invalid_argument (tensor/tensor.hpp:153-156 in the
// reference).  Calls on one host thread share one device context, created on
// first use on the device named by XTS_DEVICE (default 0), or the one given
// to Fermi::set_xts_context.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "xts_wcs.hpp"
#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/fits/fits.hxx"
#include "xtsrcmaps/irf/irf.hxx"
#include "xtsrcmaps/model_map/mm_cubature/cubature.hpp"
#include "xtsrcmaps/model_map/model_map.hxx"
#include "xtsrcmaps/psf/psf.hxx"
#include "xtsrcmaps/sky_geom/sky_geom.hxx"
#include "xtsrcmaps/xts_context.hpp"

namespace Fermi {

namespace {
thread_local xts_ctx* t_ctx = nullptr;
thread_local int32_t t_mm_flags = XTS_LAYOUT_SHWE | XTS_OUT_F64 | XTS_COMPUTE_F32;

void check(int rc, const char* what) {
    if (rc != XTS_OK) throw std::runtime_error(std::string("xtsrcmaps_amd ") + what + ": " +
                                               xts_last_error());
}

void require(bool ok, const char* what) {
    if (!ok) throw std::invalid_argument(what);
}

xts_wcs image_wcs(SkyGeom<double> const& g, size_t Nh, size_t Nw) {
    xts_wcs w = g.c_wcs();
    w.naxis1 = (int64_t)Nh;
    w.naxis2 = (int64_t)Nw;
    return w;
}
}  // namespace

xts_ctx* xts_context() {
    if (!t_ctx) {
        const char* env = std::getenv("XTS_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        check(xts_ctx_create(dev, &t_ctx), "context");
    }
    return t_ctx;
}

void set_xts_context(xts_ctx* ctx) { t_ctx = ctx; }

void set_model_map_flags(int32_t flags) {
    require((flags & ~(XTS_COMPUTE_F64)) == 0,
            "set_model_map_flags: only XTS_COMPUTE_F32 / XTS_COMPUTE_F64 apply to the "
            "[S, H, W, E] double maps of this API");
    t_mm_flags = XTS_LAYOUT_SHWE | XTS_OUT_F64 | flags;
}

// ------------------------------------------------------------------ SkyGeom
template <typename T>
SkyGeom<T>::SkyGeom(long naxis1, long naxis2, std::string const& proj_name, bool is_galactic,
                    double crpix1, double crpix2, double crval1, double crval2, double cdelt1,
                    double cdelt2, double axis_rot)
    : naxes_{naxis1, naxis2},
      crval_{crval1, crval2, 0.0},
      crpix_{crpix1, crpix2, 0.0},
      cdelt_{cdelt1, cdelt2, 0.0},
      proj_name_(proj_name) {
    int proj;
    if (proj_name == "CAR") proj = XTS_PROJ_CAR;
    else if (proj_name == "AIT") proj = XTS_PROJ_AIT;
    else throw std::invalid_argument("SkyGeom: projection '" + proj_name +
                                     "' is not supported (CAR, AIT)");
    wcs_ = xts_wcs{naxis1, naxis2, crpix1, crpix2, crval1, crval2, cdelt1, cdelt2, axis_rot,
                   proj, is_galactic ? 1 : 0};
}

template <typename T>
auto SkyGeom<T>::sph2pix(std::array<T, 2> const& ss) const -> std::array<T, 2> {
    const xts::Wcs w = xts::wcs_setup(wcs_);
    double p1, p2;
    xts::wcs_sph2pix(w, ss[0], ss[1], &p1, &p2);
    return {static_cast<T>(p1), static_cast<T>(p2)};
}

template <typename T>
auto SkyGeom<T>::sph2pix(std::pair<T, T> const& ss) const -> std::pair<T, T> {
    const auto p = sph2pix(std::array<T, 2>{ss.first, ss.second});
    return {p[0], p[1]};
}

template <typename T>
auto SkyGeom<T>::pix2sph(std::array<T, 2> const& px) const -> std::array<T, 2> {
    const xts::Wcs w = xts::wcs_setup(wcs_);
    double lon, lat;
    xts::wcs_pix2sph(w, px[0], px[1], &lon, &lat);
    return {static_cast<T>(lon), static_cast<T>(lat)};
}

template <typename T>
auto SkyGeom<T>::pix2sph(T first, T second) const -> std::array<T, 2> {
    return pix2sph(std::array<T, 2>{first, second});
}

template <typename T>
auto SkyGeom<T>::dir2sph(std::array<T, 3> const& dir) const -> std::array<T, 2> {
    T ra = std::atan2(dir[1], dir[0]) * T(xts::kR2D);
    while (ra < 0) ra += 360.;
    while (ra > 360) ra -= 360.;
    return {ra, std::asin(dir[2]) * T(xts::kR2D)};
}

template <typename T>
auto SkyGeom<T>::pix2dir(std::array<T, 2> const& px) const -> std::array<T, 3> {
    const auto s = pix2sph(px);
    return sph2dir(std::pair<T, T>{s[0], s[1]});
}

template <typename T>
auto SkyGeom<T>::sph2dir(std::pair<T, T> const& s) const -> std::array<T, 3> {
    double d[3];
    xts::sph2dir(s.first, s.second, d);
    return {static_cast<T>(d[0]), static_cast<T>(d[1]), static_cast<T>(d[2])};
}

template <typename T>
auto SkyGeom<T>::wcs_config() const -> WcsConfig {
    return WcsConfig{{naxes_[0], naxes_[1], 1},
                     crpix_,
                     crval_,
                     cdelt_,
                     wcs_.crota2,
                     proj_name_,
                     wcs_.is_galactic != 0};
}

template class SkyGeom<double>;

// ------------------------------------------------------------------ Irf
namespace Irf {

auto livetime_efficiency_factors(std::vector<double> const& logEs, IrfEffic const& effic)
    -> Tensor<double, 2> {
    Tensor<double, 2> out(2, logEs.size());
    check(xts_livetime_efficiency_factors(logEs.data(), (int64_t)logEs.size(), effic.p0.data(),
                                          effic.p1.data(), out.data()),
          "livetime_efficiency_factors");
    return out;
}

auto aeff_value(std::vector<double> const& costhet, std::vector<double> const& logEs,
                IrfData3 const& AeffData) -> Tensor<double, 2> {
    require(AeffData.params.extent(0) == AeffData.cosths.extent(0) &&
                AeffData.params.extent(1) == AeffData.logEs.extent(0),
            "aeff_value: grid shape mismatch");
    // the first parameter block of [Mc, Me, Ng] is the effective area
    const size_t Mc = AeffData.cosths.extent(0), Me = AeffData.logEs.extent(0);
    const size_t Ng = AeffData.params.extent(2);
    std::vector<double> p(Mc * Me);
    for (size_t k = 0; k < Mc * Me; ++k) p[k] = AeffData.params.data()[k * Ng];
    const xts_aeff_grid g{p.data(), AeffData.cosths.data(), AeffData.logEs.data(), (int64_t)Mc,
                          (int64_t)Me, AeffData.minCosTheta};
    Tensor<double, 2> out(costhet.size(), logEs.size());
    check(xts_aeff_value(xts_context(), costhet.data(), (int64_t)costhet.size(), logEs.data(),
                         (int64_t)logEs.size(), &g, out.data()),
          "aeff_value");
    return out;
}

void normalize_rpsf(psf::Data& psfdata) {
    auto& r = psfdata.rpsf;
    require(r.params.extent(2) == 6, "normalize_rpsf: King grids need 6 parameters");
    check(xts_normalize_rpsf(xts_context(), r.params.data(), r.logEs.data(),
                             (int64_t)r.cosths.extent(0), (int64_t)r.logEs.extent(0),
                             psfdata.psf_scaling_params.scale0, psfdata.psf_scaling_params.scale1,
                             psfdata.psf_scaling_params.scale_index),
          "normalize_rpsf");
}

// irf/irf.prepare_grid.cpp:26-102: bin centres padded by cos theta -1 / +1
// and log10 E 0 / 10, parameter blocks padded by edge duplication.
auto prepare_grid(std::vector<float> const& e_lo, std::vector<float> const& e_hi,
                  std::vector<float> const& c_lo, std::vector<float> const& c_hi,
                  std::vector<float> const& params, std::size_t n_grids) -> IrfData3 {
    const size_t Me0 = e_lo.size(), Mc0 = c_lo.size();
    require(Me0 > 1 && e_hi.size() == Me0 && Mc0 > 1 && c_hi.size() == Mc0 &&
                params.size() == n_grids * Mc0 * Me0,
            "prepare_grid: column sizes do not match");
    const size_t Me = Me0 + 2, Mc = Mc0 + 2;
    IrfData3 g{Tensor<double, 1>(Mc), Tensor<double, 1>(Me), Tensor<double, 3>(Mc, Me, n_grids),
               (double)c_lo[0]};
    g.cosths[0] = -1.0;
    g.cosths[Mc - 1] = 1.0;
    // the sum, product and log10 in float, as the reference forms them from
    // its float columns
    for (size_t c = 0; c < Mc0; ++c) {
        const float sum = c_lo[c] + c_hi[c];
        g.cosths[c + 1] = 0.5 * (double)sum;
    }
    g.logEs[0] = 0.0;
    g.logEs[Me - 1] = 10.0;
    for (size_t e = 0; e < Me0; ++e) {
        const float prod = e_lo[e] * e_hi[e];
        g.logEs[e + 1] = 0.5 * (double)std::log10(prod);
    }
    for (size_t p = 0; p < n_grids; ++p)
        for (size_t c = 0; c < Mc; ++c) {
            const size_t cs = c == 0 ? 0 : (c >= Mc0 ? Mc0 - 1 : c - 1);
            for (size_t e = 0; e < Me; ++e) {
                const size_t es = e == 0 ? 0 : (e >= Me0 ? Me0 - 1 : e - 1);
                g.params[c, e, p] = params[(p * Mc0 + cs) * Me0 + es];
            }
        }
    return g;
}

auto prepare_grid(fits::TablePars const& pars) -> IrfData3 {
    const auto& x = pars.extents;
    require(x.size() >= 5 && x[0] > 1 && x[1] == x[0] && x[2] > 1 && x[3] == x[2] &&
                pars.rowdata.extent(0) == 1,
            "prepare_grid: not an IRF grid table");
    for (size_t k = 4; k < x.size(); ++k)
        require(x[k] == x[0] * x[2], "prepare_grid: parameter column size");
    auto col = [&](size_t k) {
        const float* p = &pars.rowdata[0, pars.offsets[k]];
        return std::vector<float>(p, p + x[k]);
    };
    std::vector<float> params;
    for (size_t k = 4; k < x.size(); ++k) {
        const auto c = col(k);
        params.insert(params.end(), c.begin(), c.end());
    }
    return prepare_grid(col(0), col(1), col(2), col(3), params, x.size() - 4);
}

namespace {
// irf.prepare_scale.cpp:7-12 and irf.prepare_effic.cpp:8-21
IrfScale prepare_scale(fits::TablePars const& p) {
    require(p.rowdata.extent(0) == 1 && p.rowdata.extent(1) == 3, "PSF scaling table shape");
    return {p.rowdata[0, 0], p.rowdata[0, 1], p.rowdata[0, 2]};
}
IrfEffic prepare_effic(fits::TablePars const& p) {
    require(p.rowdata.extent(0) == 2 && p.rowdata.extent(1) == 6, "efficiency table shape");
    IrfEffic e{};
    for (size_t k = 0; k < 6; ++k) {
        e.p0[k] = p.rowdata[0, k];
        e.p1[k] = p.rowdata[1, k];
    }
    return e;
}
}  // namespace

auto load_aeff(std::string const& filename) -> std::optional<aeff::Pass8FB> {
    const char* names[6] = {"EFFECTIVE AREA_FRONT", "PHI_DEPENDENCE_FRONT",
                            "EFFICIENCY_PARAMS_FRONT", "EFFECTIVE AREA_BACK",
                            "PHI_DEPENDENCE_BACK", "EFFICIENCY_PARAMS_BACK"};
    std::optional<fits::TablePars> t[6];
    for (int k = 0; k < 6; ++k)
        if (!(t[k] = fits::read_irf_pars(filename, names[k]))) return std::nullopt;
    return aeff::Pass8FB{{prepare_grid(*t[0]), prepare_grid(*t[1]), prepare_effic(*t[2])},
                         {prepare_grid(*t[3]), prepare_grid(*t[4]), prepare_effic(*t[5])}};
}

auto load_psf(std::string const& filename) -> std::optional<psf::Pass8FB> {
    const char* names[6] = {"RPSF_FRONT", "PSF_SCALING_PARAMS_FRONT", "FISHEYE_CORRECTION_FRONT",
                            "RPSF_BACK",  "PSF_SCALING_PARAMS_BACK",  "FISHEYE_CORRECTION_BACK"};
    std::optional<fits::TablePars> t[6];
    for (int k = 0; k < 6; ++k)
        if (!(t[k] = fits::read_irf_pars(filename, names[k]))) return std::nullopt;
    psf::Pass8FB out{{prepare_grid(*t[0]), prepare_scale(*t[1]), prepare_grid(*t[2])},
                     {prepare_grid(*t[3]), prepare_scale(*t[4]), prepare_grid(*t[5])}};
    normalize_rpsf(out.front);
    normalize_rpsf(out.back);
    return out;
}

}  // namespace Irf

// ------------------------------------------------------------------ Exposure
namespace Exposure {

auto map(Obs::ExposureCubeData const& data) -> ExposureMap {
    const size_t npix = 12 * (size_t)data.nside * data.nside;
    require(data.cosbins.size() == npix * data.nbrbins, "Exposure::map: cube size mismatch");
    Tensor<double, 2> p(npix, (size_t)data.nbrbins);
    std::copy(data.cosbins.begin(), data.cosbins.end(), p.data());
    return {data.nside, data.nbrbins, p};
}

auto costhetas(Obs::ExposureCubeData const& data) -> std::vector<double> {
    std::vector<double> v(data.nbrbins);
    check(xts_costhetas((int64_t)data.nbrbins, data.cosmin, data.thetabin ? 1 : 0, v.data()),
          "costhetas");
    return v;
}

auto contract(Tensor<double, 2> const& A, Tensor<double, 2> const& B) -> Tensor<double, 2> {
    require(A.extent(1) == B.extent(0), "contract: inner extents differ");
    Tensor<double, 2> out(A.extent(0), B.extent(1));
    check(xts_contract(xts_context(), A.data(), B.data(), (int64_t)A.extent(0),
                       (int64_t)A.extent(1), (int64_t)B.extent(1), out.data(), XTS_PREC_FP64),
          "contract");
    return out;
}

auto src_cosbins(Tensor<double, 2> const& src_sph, ExposureMap const& expmap)
    -> Tensor<double, 2> {
    require(src_sph.extent(1) == 2, "src_cosbins: src_sph must be [S, 2]");
    Tensor<double, 2> out(src_sph.extent(0), expmap.nbins);
    check(xts_src_cosbins(xts_context(), src_sph.data(), (int64_t)src_sph.extent(0),
                          expmap.params.data(), (int64_t)expmap.nside, (int64_t)expmap.nbins,
                          out.data()),
          "src_cosbins");
    return out;
}

auto exposure(Tensor<double, 2> const& ecb, Tensor<double, 2> const& wecb,
              Tensor<double, 2> const& front_aeff, Tensor<double, 2> const& back_aeff,
              Tensor<double, 2> const& front_ltfs) -> Tensor<double, 2> {
    const size_t S = ecb.extent(0), C = ecb.extent(1), E = front_aeff.extent(1);
    require(wecb.extent(0) == S && wecb.extent(1) == C && front_aeff.extent(0) == C &&
                back_aeff.extent(0) == C && back_aeff.extent(1) == E &&
                front_ltfs.extent(0) == 2 && front_ltfs.extent(1) == E,
            "exposure: shape mismatch");
    Tensor<double, 2> out(S, E);
    check(xts_exposure(xts_context(), ecb.data(), wecb.data(), front_aeff.data(),
                       back_aeff.data(), front_ltfs.data(), (int64_t)S, (int64_t)C, (int64_t)E,
                       out.data(), XTS_PREC_FP64),
          "exposure");
    return out;
}

}  // namespace Exposure

// ------------------------------------------------------------------ Psf
namespace Psf {

namespace {
xts_psf_grid psf_grid(Irf::IrfData3 const& d) {
    return {d.params.data(), d.cosths.data(), d.logEs.data(), (int64_t)d.cosths.extent(0),
            (int64_t)d.logEs.extent(0)};
}
}  // namespace

// psf/psf.fast_separation_lower_index.cxx:5-22 (host integer arithmetic)
auto fast_separation_lower_index(Tensor<double, 1> seps) -> Tensor<int, 1> {
    const double recipstep = 399.0 / std::log(7e5);
    Tensor<int, 1> index(seps.extent(0));
    std::transform(seps.begin(), seps.end(), index.begin(), [recipstep](double v) {
        return v < 1. ? 1 : static_cast<int>(std::floor(1. + std::log(1e4 * v) * recipstep));
    });
    return index;
}

auto king(Irf::psf::Data const& data) -> Tensor<double, 3> {
    const auto g = psf_grid(data.rpsf);
    Tensor<double, 3> out(sep_arr_len, (size_t)g.Mc, (size_t)g.Me);
    check(xts_king(xts_context(), &g, out.data()), "king");
    return out;
}

auto bilerp(std::vector<double> const& costhetas, std::vector<double> const& logEs,
            Tensor<double, 1> const& par_cosths, Tensor<double, 1> const& par_logEs,
            Tensor<double, 3> const& kings) -> Tensor<double, 3> {
    require(kings.extent(0) == sep_arr_len && kings.extent(1) == par_cosths.extent(0) &&
                kings.extent(2) == par_logEs.extent(0),
            "bilerp: kings must be [401, Mc, Me]");
    const xts_psf_grid g{nullptr, par_cosths.data(), par_logEs.data(),
                         (int64_t)par_cosths.extent(0), (int64_t)par_logEs.extent(0)};
    Tensor<double, 3> out(costhetas.size(), sep_arr_len, logEs.size());
    check(xts_bilerp(xts_context(), costhetas.data(), (int64_t)costhetas.size(), logEs.data(),
                     (int64_t)logEs.size(), &g, kings.data(), out.data()),
          "bilerp");
    return out;
}

auto corrected_exposure_psf(Tensor<double, 3> const& obs_psf, Tensor<double, 2> const& obs_aeff,
                            Tensor<double, 2> const& ecb, Tensor<double, 2> const& wecb,
                            Tensor<double, 2> const& front_LTF) -> Tensor<double, 3> {
    const size_t C = obs_psf.extent(0), E = obs_psf.extent(2), S = ecb.extent(0);
    require(obs_psf.extent(1) == sep_arr_len && obs_aeff.extent(0) == C &&
                obs_aeff.extent(1) == E && ecb.extent(1) == C && wecb.extent(0) == S &&
                wecb.extent(1) == C && front_LTF.extent(1) == E,
            "corrected_exposure_psf: shape mismatch");
    Tensor<double, 3> out(S, sep_arr_len, E);
    check(xts_corrected_exposure_psf(xts_context(), obs_psf.data(), obs_aeff.data(), ecb.data(),
                                     wecb.data(), front_LTF.data(), (int64_t)S, (int64_t)C,
                                     (int64_t)E, out.data()),
          "corrected_exposure_psf");
    return out;
}

auto mean_psf(Tensor<double, 3> const& front_corrected_psf,
              Tensor<double, 3> const& back_corrected_psf, Tensor<double, 2> const& exposures)
    -> Tensor<double, 3> {
    const size_t S = front_corrected_psf.extent(0), E = front_corrected_psf.extent(2);
    require(front_corrected_psf.extents() == back_corrected_psf.extents() &&
                exposures.extent(0) == S && exposures.extent(1) == E,
            "mean_psf: shape mismatch");
    Tensor<double, 3> out(S, sep_arr_len, E);
    check(xts_mean_psf(xts_context(), front_corrected_psf.data(), back_corrected_psf.data(),
                       exposures.data(), (int64_t)S, (int64_t)E, out.data()),
          "mean_psf");
    return out;
}

auto partial_total_integral(Tensor<double, 3> const& mean_psf)
    -> std::pair<Tensor<double, 3>, Tensor<double, 2>> {
    const size_t S = mean_psf.extent(0), E = mean_psf.extent(2);
    require(mean_psf.extent(1) == sep_arr_len, "partial_total_integral: [S, 401, E] expected");
    Tensor<double, 3> scratch(mean_psf.begin(), mean_psf.end(), mean_psf.extents());
    Tensor<double, 3> par(S, sep_arr_len, E);
    Tensor<double, 2> tot(S, E);
    check(xts_partial_total_integral(xts_context(), scratch.data(), (int64_t)S, (int64_t)E,
                                     par.data(), tot.data()),
          "partial_total_integral");
    return {par, tot};
}

auto normalize(Tensor<double, 3>& mean_psf, Tensor<double, 2> const& total_integrals) -> void {
    const size_t S = mean_psf.extent(0), E = mean_psf.extent(2);
    require(total_integrals.extent(0) == S && total_integrals.extent(1) == E,
            "normalize: shape mismatch");
    check(xts_normalize(xts_context(), mean_psf.data(), total_integrals.data(), (int64_t)S,
                        (int64_t)E),
          "normalize");
}

auto psf_lookup_table_and_partial_integrals(
    Irf::psf::Pass8FB const& data, std::vector<double> const& costhetas,
    std::vector<double> const& logEs, Tensor<double, 2> const& front_aeff,
    Tensor<double, 2> const& back_aeff, Tensor<double, 2> const& ecb,
    Tensor<double, 2> const& wecb, Tensor<double, 2> const& front_LTF,
    Tensor<double, 2> const& exposure) -> std::tuple<Tensor<double, 3>, Tensor<double, 3>> {
    const size_t S = exposure.extent(0), E = exposure.extent(1), C = costhetas.size();
    require(logEs.size() == E && front_aeff.extent(0) == C && front_aeff.extent(1) == E &&
                back_aeff.extent(0) == C && ecb.extent(0) == S && ecb.extent(1) == C &&
                wecb.extent(0) == S && front_LTF.extent(1) == E,
            "psf_lookup_table_and_partial_integrals: shape mismatch");
    const auto gf = psf_grid(data.front.rpsf), gb = psf_grid(data.back.rpsf);
    Tensor<double, 3> upsf(S, sep_arr_len, E), par(S, sep_arr_len, E);
    check(xts_psf_lookup_table_and_partial_integrals(
              xts_context(), &gf, &gb, costhetas.data(), (int64_t)C, logEs.data(), (int64_t)E,
              front_aeff.data(), back_aeff.data(), ecb.data(), wecb.data(), front_LTF.data(),
              exposure.data(), (int64_t)S, upsf.data(), par.data(), XTS_PREC_FP64),
          "psf_lookup_table_and_partial_integrals");
    return {upsf, par};
}

auto compute_psf_data(Obs::XtObs const& obs, Irf::XtIrf const& irf, Exposure::XtExp const& exp)
    -> XtPsf {
    auto [upsf, par] = psf_lookup_table_and_partial_integrals(
        irf.psf_irf, exp.exp_costhetas, obs.logEs, exp.front_aeff, exp.back_aeff,
        exp.src_exposure_cosbins, exp.src_weighted_exposure_cosbins, irf.front_LTF,
        exp.exposure);
    return {upsf, par};
}

auto make_kernel(size_t const Nk, SkyGeom<double> const& skygeom,
                 Tensor<double, 3> const& uPsf) -> Tensor<double, 3> {
    require(Nk % 2 == 1 && uPsf.extent(0) >= 1 && uPsf.extent(1) == sep_arr_len,
            "make_kernel: Nk must be odd and uPsf [S, 401, E]");
    const size_t E = uPsf.extent(2);
    const xts_wcs w = skygeom.c_wcs();
    Tensor<double, 3> out(E, Nk, Nk);
    // uPsf[0, :, :] is the first sep_arr_len x E block
    check(xts_psf_make_kernel(xts_context(), (int64_t)Nk, &w, w.crpix1, w.crpix2, uPsf.data(),
                              (int64_t)E, out.data()),
          "make_kernel");
    return out;
}

}  // namespace Psf

// ------------------------------------------------------------------ cubature
auto cubature_convolve_map_psf(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sph,
                               Tensor<double, 3> const& psf_lut, SkyGeom<double> const& skygeom,
                               std::vector<double> const& rule_uvw) -> Tensor<double, 4> {
    require(src_sph.extent(1) == 2 && psf_lut.extent(0) == src_sph.extent(0) &&
                psf_lut.extent(1) == Psf::sep_arr_len && rule_uvw.size() % 3 == 0,
            "cubature_convolve_map_psf: shape mismatch");
    const size_t S = src_sph.extent(0), E = psf_lut.extent(2);
    const xts_wcs w = image_wcs(skygeom, Nh, Nw);
    Tensor<double, 4> out(S, Nh, Nw, E);
    check(xts_cubature_convolve_map_psf(xts_context(), &w, src_sph.data(), (int64_t)S,
                                        psf_lut.data(), (int64_t)E,
                                        rule_uvw.empty() ? nullptr : rule_uvw.data(),
                                        (int64_t)(rule_uvw.size() / 3), t_mm_flags, out.data()),
          "cubature_convolve_map_psf");
    return out;
}

auto cubature_convolve_map_psf(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sph,
                               Tensor<double, 3> const& psf_lut, SkyGeom<double> const& skygeom)
    -> Tensor<double, 4> {
    return cubature_convolve_map_psf(Nh, Nw, src_sph, psf_lut, skygeom, {});
}

// ------------------------------------------------------------------ ModelMap
namespace ModelMap {

auto point_src_model_map_wcs(size_t Nh, size_t Nw, SkyGeom<double> const& skygeom,
                             Tensor<double, 2> const& src_sph,
                             std::vector<std::string> const& src_names,
                             Tensor<double, 2> const& exposures, Tensor<double, 3> const& uPsf,
                             Tensor<double, 3> const& partial_integrals) -> Tensor<double, 4> {
    const size_t S = src_sph.extent(0), E = exposures.extent(1);
    require(src_sph.extent(1) == 2 && exposures.extent(0) == S && uPsf.extent(0) == S &&
                uPsf.extent(1) == Psf::sep_arr_len && uPsf.extent(2) == E &&
                partial_integrals.extents() == uPsf.extents() &&
                (src_names.empty() || src_names.size() == S),
            "point_src_model_map_wcs: shape mismatch");
    const xts_wcs w = image_wcs(skygeom, Nh, Nw);
    Tensor<double, 4> out(S, Nh, Nw, E);
    check(xts_point_src_model_map_wcs(xts_context(), &w, src_sph.data(), (int64_t)S,
                                      exposures.data(), uPsf.data(), partial_integrals.data(),
                                      (int64_t)E, nullptr, 0, t_mm_flags, out.data()),
          "point_src_model_map_wcs");
    return out;
}

void scale_map_by_solid_angle(Tensor<double, 4>& model_map, SkyGeom<double> const& skygeom) {
    const size_t S = model_map.extent(0), H = model_map.extent(1), W = model_map.extent(2),
                 E = model_map.extent(3);
    const xts_wcs w = image_wcs(skygeom, H, W);
    std::vector<double> omega(H * W);
    check(xts_solid_angle(xts_context(), &w, omega.data()), "solid_angle");
    check(xts_scale_map(xts_context(), model_map.data(), (int64_t)S, (int64_t)H, (int64_t)W,
                        (int64_t)E, omega.data(), XTS_SCALE_PIXEL),
          "scale_map_by_solid_angle");
}

void scale_map_by_exposure(Tensor<double, 4>& model_map, Tensor<double, 2> const& exposures) {
    require(exposures.extent(0) == model_map.extent(0) &&
                exposures.extent(1) == model_map.extent(3),
            "scale_map_by_exposure: shape mismatch");
    check(xts_scale_map(xts_context(), model_map.data(), (int64_t)model_map.extent(0),
                        (int64_t)model_map.extent(1), (int64_t)model_map.extent(2),
                        (int64_t)model_map.extent(3), exposures.data(), XTS_SCALE_SRC_ENERGY),
          "scale_map_by_exposure");
}

void scale_map_by_correction_factors(Tensor<double, 4>& model_map,
                                     Tensor<double, 2> const& factor) {
    require(factor.extent(0) == model_map.extent(0) && factor.extent(1) == model_map.extent(3),
            "scale_map_by_correction_factors: shape mismatch");
    check(xts_scale_map(xts_context(), model_map.data(), (int64_t)model_map.extent(0),
                        (int64_t)model_map.extent(1), (int64_t)model_map.extent(2),
                        (int64_t)model_map.extent(3), factor.data(), XTS_SCALE_SRC_ENERGY),
          "scale_map_by_correction_factors");
}

auto map_correction_factor(Tensor<double, 2> const& inv_map_integ,
                           Tensor<double, 1> const& psf_radius,
                           std::vector<bool> const& is_in_fov, Tensor<double, 3> const& mean_psf,
                           Tensor<double, 3> const& partial_integrals) -> Tensor<double, 2> {
    const size_t S = inv_map_integ.extent(0), E = inv_map_integ.extent(1);
    require(psf_radius.extent(0) == S && is_in_fov.size() == S && mean_psf.extent(0) == S &&
                mean_psf.extent(2) == E && partial_integrals.extents() == mean_psf.extents(),
            "map_correction_factor: shape mismatch");
    std::vector<uint8_t> fov(S);
    for (size_t s = 0; s < S; ++s) fov[s] = is_in_fov[s] ? 1 : 0;
    Tensor<double, 2> out(S, E);
    check(xts_map_correction_factor(xts_context(), inv_map_integ.data(), psf_radius.data(),
                                    fov.data(), mean_psf.data(), partial_integrals.data(),
                                    (int64_t)S, (int64_t)E, out.data()),
          "map_correction_factor");
    return out;
}

auto psf_boundary_radius(size_t Nh, size_t Nw, Tensor<double, 2> const& src_sph,
                         SkyGeom<double> const& skygeom)
    -> std::pair<Tensor<double, 1>, std::vector<bool>> {
    require(src_sph.extent(1) == 2, "psf_boundary_radius: src_sph must be [S, 2]");
    const size_t S = src_sph.extent(0);
    const xts_wcs w = image_wcs(skygeom, Nh, Nw);
    Tensor<double, 1> radius(S);
    std::vector<uint8_t> fov(S);
    check(xts_psf_boundary_radius(xts_context(), &w, src_sph.data(), (int64_t)S, radius.data(),
                                  fov.data()),
          "psf_boundary_radius");
    return {radius, std::vector<bool>(fov.begin(), fov.end())};
}

auto map_integral(Tensor<double, 4> const& model_map, Tensor<double, 2> const& src_sph,
                  SkyGeom<double> const& skygeom, Tensor<double, 1> const& psf_radius,
                  std::vector<bool> const& is_in_fov) -> Tensor<double, 2> {
    const size_t S = model_map.extent(0), H = model_map.extent(1), W = model_map.extent(2),
                 E = model_map.extent(3);
    require(src_sph.extent(0) == S && psf_radius.extent(0) == S && is_in_fov.size() == S,
            "map_integral: shape mismatch");
    const xts_wcs w = image_wcs(skygeom, H, W);
    std::vector<uint8_t> fov(S);
    for (size_t s = 0; s < S; ++s) fov[s] = is_in_fov[s] ? 1 : 0;
    Tensor<double, 2> out(S, E);
    check(xts_map_integral(xts_context(), &w, model_map.data(), src_sph.data(), (int64_t)S,
                           (int64_t)E, psf_radius.data(), fov.data(), out.data()),
          "map_integral");
    return out;
}

auto diffuse_src_model_map_wcs(size_t const Nh, size_t const Nw, SkyGeom<double> const& skygeom,
                               Tensor<double, 2> const& src_sph,
                               std::vector<std::string> const& src_names,
                               Tensor<double, 2> const& exposures, Tensor<double, 3> const& uPsf,
                               Tensor<double, 3> const& partial_integrals,
                               double const target_pix_size) -> Tensor<double, 4> {
    require(!src_names.empty() && exposures.extent(0) >= 1 && uPsf.extent(0) >= 1 &&
                uPsf.extent(1) == Psf::sep_arr_len && uPsf.extent(2) == exposures.extent(1),
            "diffuse_src_model_map_wcs: a MapCube file name, exposures [S, E], uPsf [S, 401, E]");
    (void)src_sph;            // the kernel is centred on skygeom's reference direction
    (void)partial_integrals;  // no point-source correction factor for a diffuse source
    const size_t E = exposures.extent(1);
    xts_wcs cube_wcs{};
    int64_t planes = 0;
    check(xts_read_mapcube(src_names[0].c_str(), &cube_wcs, &planes, nullptr, nullptr, nullptr),
          "read_mapcube");
    std::vector<float> cube((size_t)planes * cube_wcs.naxis1 * cube_wcs.naxis2);
    check(xts_read_mapcube(src_names[0].c_str(), &cube_wcs, &planes, cube.data(), nullptr,
                           nullptr),
          "read_mapcube");
    const xts_wcs w = image_wcs(skygeom, Nh, Nw);
    std::vector<double> ehw(E * Nh * Nw);
    // row 0 of exposures and uPsf: the diffuse source's
    check(xts_diffuse_src_model_map_wcs(xts_context(), &w, &cube_wcs, cube.data(), planes,
                                        uPsf.data(), exposures.data(), (int64_t)E,
                                        target_pix_size, 0.0, ehw.data()),
          "diffuse_src_model_map_wcs");
    Tensor<double, 4> out(1, Nh, Nw, E);
    for (size_t e = 0; e < E; ++e)
        for (size_t h = 0; h < Nh; ++h)
            for (size_t x = 0; x < Nw; ++x) out[0, h, x, e] = ehw[(e * Nh + h) * Nw + x];
    return out;
}

}  // namespace ModelMap

namespace FFT {
auto convolve_map_psf(Tensor<double, 3> const& image, Tensor<double, 3> const& kernel)
    -> Tensor<double, 3> {
    const size_t C = image.extent(0), H = image.extent(1), W = image.extent(2);
    const size_t KH = kernel.extent(1), KW = kernel.extent(2);
    require(kernel.extent(0) == C, "convolve_map_psf: channel dimensions of image and kernel "
                                   "must match");
    Tensor<double, 3> out(C, H + KH - 1, W + KW - 1);
    check(xts_fft_convolve_map_psf(xts_context(), image.data(), (int64_t)C, (int64_t)H,
                                   (int64_t)W, kernel.data(), (int64_t)KH, (int64_t)KW,
                                   out.data()),
          "convolve_map_psf");
    return out;
}
}  // namespace FFT

// ---------------------------------------------------------------- fits
namespace fits {

auto read_irf_pars(std::string const& filename, std::string tblname)
    -> std::optional<TablePars> {
    int64_t nc = 0, nr = 0, rw = 0;
    if (xts_read_irf_pars(filename.c_str(), tblname.c_str(), &nc, nullptr, 0, &nr, &rw,
                          nullptr) != XTS_OK) {
        std::fprintf(stderr, "read_irf_pars: %s\n", xts_last_error());
        return std::nullopt;
    }
    std::vector<int64_t> ext(nc > 0 ? nc : 1);
    TablePars t{{}, {}, Tensor<float, 2>((size_t)nr, (size_t)rw)};
    if (xts_read_irf_pars(filename.c_str(), tblname.c_str(), &nc, ext.data(), nc, &nr, &rw,
                          t.rowdata.data()) != XTS_OK) {
        std::fprintf(stderr, "read_irf_pars: %s\n", xts_last_error());
        return std::nullopt;
    }
    size_t off = 0;
    for (int64_t c = 0; c < nc; ++c) {
        t.extents.push_back((size_t)ext[c]);
        t.offsets.push_back(off);
        off += (size_t)ext[c];
    }
    return t;
}

auto read_expcube(std::string const& filename, std::string const& tblname)
    -> std::optional<Obs::ExposureCubeData> {
    xts_expcube_info info{};
    if (xts_read_expcube(filename.c_str(), tblname.c_str(), &info, nullptr, nullptr, nullptr) !=
        XTS_OK) {
        std::fprintf(stderr, "read_expcube: %s\n", xts_last_error());
        return std::nullopt;
    }
    const size_t npix = 12 * (size_t)info.nside * info.nside;
    Obs::ExposureCubeData d;
    d.cosbins.resize(npix * info.nbrbins);
    d.ra.resize(npix);
    d.dec.resize(npix);
    if (xts_read_expcube(filename.c_str(), tblname.c_str(), &info, d.cosbins.data(),
                         d.ra.data(), d.dec.data()) != XTS_OK) {
        std::fprintf(stderr, "read_expcube: %s\n", xts_last_error());
        return std::nullopt;
    }
    d.nside = (unsigned)info.nside;
    d.nbrbins = (unsigned)info.nbrbins;
    d.cosmin = info.cosmin;
    d.ordering = info.ordering;
    d.coordsys = info.coordsys;
    d.thetabin = info.thetabin != 0;
    return d;
}

auto write_src_model(std::string const& filename, Tensor<double, 4>& model_map,
                     std::vector<std::string> const& srcs) -> void {
    const size_t S = model_map.extent(0);
    require(srcs.size() == S, "write_src_model: one name per source");
    std::vector<const char*> names(S);
    for (size_t s = 0; s < S; ++s) names[s] = srcs[s].c_str();
    check(xts_write_src_model(filename.c_str(), model_map.data(), (int64_t)S,
                              (int64_t)model_map.extent(1), (int64_t)model_map.extent(2),
                              (int64_t)model_map.extent(3), names.data(),
                              XTS_LAYOUT_SHWE | XTS_OUT_F64, nullptr),
          "write_src_model");
}

}  // namespace fits
}  // namespace Fermi
