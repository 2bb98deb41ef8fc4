// fermi_pipeline.cpp -- the reference's model-map pipeline (xtsrcmaps/main.cxx:
// 17-40: compute_exposure, compute_psf_data, compute_srcmaps) run through this
// library's Fermi:: C++ API, exactly as a caller of the reference would, plus
// the stage-by-stage forms of the mean PSF (psf.hxx:50-84) and of the model map
// (model_map.point_src_model_map_wcs.cxx:6-45).
//
//   fermi_pipeline <inputs.xtsa> <outputs.xtsa> [<srcmaps.fits>]
//   fermi_pipeline --fits <aeff.fits> <psf.fits> <ltcube.fits> <inputs.xtsa>
//                  <outputs.xtsa> [<srcmaps.fits>]
//
// With --fits the IRFs come from Irf::load_aeff / load_psf (irf/irf.cxx:24-68)
// and the livetime cubes from fits::read_expcube (tables EXPOSURE and
// WEIGHTED_EXPOSURE), as the reference's collect_irf_data and observation
// setup read them; the inputs then hold only the sources, energies and WCS.
//
// Arrays travel in a small binary container: "XTSA", u32 count, then per
// array u32 name length, name, u32 rank, u64 extents[rank], f64 data.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "xtsrcmaps/exposure/exposure.hxx"
#include "xtsrcmaps/fits/fits.hxx"
#include "xtsrcmaps/irf/irf.hxx"
#include "xtsrcmaps/model_map/mm_cubature/cubature.hpp"
#include "xtsrcmaps/model_map/model_map.hxx"
#include "xtsrcmaps/psf/psf.hxx"

struct Array {
    std::vector<uint64_t> shape;
    std::vector<double> data;
};
using Arrays = std::map<std::string, Array>;

static Arrays read_arrays(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    char magic[4];
    f.read(magic, 4);
    if (std::string(magic, 4) != "XTSA") throw std::runtime_error("bad magic");
    uint32_t n;
    f.read((char*)&n, 4);
    Arrays out;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t len, rank;
        f.read((char*)&len, 4);
        std::string name(len, '\0');
        f.read(name.data(), len);
        f.read((char*)&rank, 4);
        Array a;
        a.shape.resize(rank);
        f.read((char*)a.shape.data(), 8 * rank);
        size_t count = 1;
        for (auto d : a.shape) count *= d;
        a.data.resize(count);
        f.read((char*)a.data.data(), 8 * count);
        out[name] = std::move(a);
    }
    if (!f) throw std::runtime_error("truncated input");
    return out;
}

static void write_arrays(const char* path, const Arrays& arrays) {
    std::ofstream f(path, std::ios::binary);
    f.write("XTSA", 4);
    const uint32_t n = (uint32_t)arrays.size();
    f.write((const char*)&n, 4);
    for (auto& [name, a] : arrays) {
        const uint32_t len = (uint32_t)name.size(), rank = (uint32_t)a.shape.size();
        f.write((const char*)&len, 4);
        f.write(name.data(), len);
        f.write((const char*)&rank, 4);
        f.write((const char*)a.shape.data(), 8 * rank);
        f.write((const char*)a.data.data(), 8 * a.data.size());
    }
}

template <typename T, std::size_t R>
static Array from_tensor(const Fermi::Tensor<T, R>& t) {
    Array a;
    for (std::size_t d = 0; d < R; ++d) a.shape.push_back(t.extent(d));
    a.data.assign(t.begin(), t.end());
    return a;
}

template <std::size_t R>
static Fermi::Tensor<double, R> tensor(const Array& a) {
    typename Fermi::Tensor<double, R>::ExtentsType ext{};
    if (a.shape.size() != R) throw std::invalid_argument("rank mismatch");
    for (std::size_t d = 0; d < R; ++d) ext[d] = a.shape[d];
    return Fermi::Tensor<double, R>(a.data, ext);
}

static Fermi::Irf::IrfData3 grid(const Arrays& in, const std::string& side, const char* kind) {
    const std::string p = std::string(kind) + "_" + side;
    Fermi::Irf::IrfData3 g{tensor<1>(in.at(p + "_cosths")), tensor<1>(in.at(p + "_logEs")),
                           tensor<3>(in.at(p + "_params")), 0.0};
    if (in.count(p + "_minc")) g.minCosTheta = in.at(p + "_minc").data[0];
    return g;
}

int main(int argc, char** argv) {
    const bool from_fits = argc > 1 && std::string(argv[1]) == "--fits";
    if (from_fits ? (argc != 7 && argc != 8) : (argc != 3 && argc != 4)) {
        std::fprintf(stderr,
                     "usage: %s <inputs.xtsa> <outputs.xtsa> [<srcmaps.fits>]\n"
                     "       %s --fits <aeff.fits> <psf.fits> <ltcube.fits> <inputs.xtsa> "
                     "<outputs.xtsa> [<srcmaps.fits>]\n",
                     argv[0], argv[0]);
        return 2;
    }
    // file arguments after the optional FITS inputs
    const int a0 = from_fits ? 5 : 1;
    const char* fits_out = argc > a0 + 2 ? argv[a0 + 2] : nullptr;
    try {
        const Arrays in = read_arrays(argv[a0]);
        const auto& w = in.at("wcs").data;  // naxis1 naxis2 crpix1 crpix2 crval1 crval2 cdelt1 cdelt2 crota2 proj
        const long Nh = (long)w[0], Nw = (long)w[1];
        const std::vector<double> logE = in.at("logE").data;

        // observation (obs_types.hxx) and sources (source.hxx)
        Fermi::Obs::XtObs obs{Nh, Nw, {}, logE, {}, {},
                              Fermi::SkyGeom<double>(Nh, Nw, w[9] == 0 ? "CAR" : "AIT", false,
                                                     w[2], w[3], w[4], w[5], w[6], w[7], w[8])};
        for (double le : logE) obs.energies.push_back(std::pow(10.0, le));
        if (from_fits) {
            auto e = Fermi::fits::read_expcube(argv[4], "EXPOSURE");
            auto we = Fermi::fits::read_expcube(argv[4], "WEIGHTED_EXPOSURE");
            if (!e || !we) throw std::runtime_error("cannot read the livetime cube");
            obs.exp_cube = std::move(*e);
            obs.weighted_exp_cube = std::move(*we);
        } else {
            const unsigned nside = (unsigned)in.at("nside").data[0];
            for (auto* c : {&obs.exp_cube, &obs.weighted_exp_cube}) {
                const auto& a = in.at(c == &obs.exp_cube ? "cube" : "wcube");
                c->cosbins.assign(a.data.begin(), a.data.end());
                c->nside = nside;
                c->nbrbins = (unsigned)a.shape[1];
                c->cosmin = 0.0;
                c->thetabin = true;
            }
        }
        // the sources as the reference's source-model reader leaves them
        // (Source::XtSrc; main.cxx passes src.point into the path)
        Fermi::Config::XtCfg cfg;
        Fermi::Source::XtSrc xsrc;
        auto& src = xsrc.point;
        src.sph_locs = tensor<2>(in.at("src_sph"));
        for (size_t s = 0; s < src.sph_locs.extent(0); ++s) {
            Fermi::Source::PointSource ps;
            ps.name = "src" + std::to_string(s);
            ps.spectrum = Fermi::Source::NormalSpectrum{{}, Fermi::Source::SpectrumType::PowerLaw};
            ps.spatial_model = Fermi::Source::SkyDirFunctionSpatialModel{};
            src.srcs.push_back(ps);
        }
        src.names = Fermi::Source::source_names(src.srcs);

        // IRFs: King sections normalised in place as the reference's loader does
        Fermi::Irf::XtIrf irf;
        if (from_fits) {
            auto a = Fermi::Irf::load_aeff(argv[2]);
            auto p = Fermi::Irf::load_psf(argv[3]);
            if (!a || !p) throw std::runtime_error("cannot read the IRF files");
            irf.aeff_irf = std::move(*a);
            irf.psf_irf = std::move(*p);
        } else {
            const auto& sc = in.at("psf_scale").data;
            for (const std::string side : {"front", "back"}) {
                auto& pd = side == "front" ? irf.psf_irf.front : irf.psf_irf.back;
                pd.rpsf = grid(in, side, "psf");
                pd.psf_scaling_params = {(float)sc[0], (float)sc[1], (float)sc[2]};
                Fermi::Irf::normalize_rpsf(pd);
                auto& ad = side == "front" ? irf.aeff_irf.front : irf.aeff_irf.back;
                ad.effective_area = grid(in, side, "aeff");
            }
            Fermi::Irf::IrfEffic effic{};
            for (int k = 0; k < 6; ++k) {
                effic.p0[k] = (float)in.at("effic_p0").data[k];
                effic.p1[k] = (float)in.at("effic_p1").data[k];
            }
            irf.aeff_irf.front.efficiency_params = effic;
        }
        irf.front_LTF = Fermi::Irf::livetime_efficiency_factors(
            obs.logEs, irf.aeff_irf.front.efficiency_params);

        // the pipeline of main.cxx:26-30
        auto exp = Fermi::Exposure::compute_exposure(cfg, obs, xsrc.point, irf);
        auto psf = Fermi::Psf::compute_psf_data(obs, irf, exp);
        auto map = Fermi::ModelMap::compute_srcmaps(obs, xsrc.point, exp, psf);
        // the output step main.cxx:40 leaves commented out
        if (fits_out) Fermi::fits::write_src_model(fits_out, map, src.names);

        // mean PSF stage by stage
        Fermi::Tensor<double, 3> corr[2];
        for (int k = 0; k < 2; ++k) {
            const auto& pd = k == 0 ? irf.psf_irf.front : irf.psf_irf.back;
            const auto kings = Fermi::Psf::king(pd);
            const auto obs_psf = Fermi::Psf::bilerp(exp.exp_costhetas, obs.logEs, pd.rpsf.cosths,
                                                    pd.rpsf.logEs, kings);
            corr[k] = Fermi::Psf::corrected_exposure_psf(
                obs_psf, k == 0 ? exp.front_aeff : exp.back_aeff, exp.src_exposure_cosbins,
                exp.src_weighted_exposure_cosbins, irf.front_LTF);
        }
        auto mean = Fermi::Psf::mean_psf(corr[0], corr[1], exp.exposure);
        const auto [par, tot] = Fermi::Psf::partial_total_integral(mean);
        Fermi::Psf::normalize(mean, tot);

        // model map stage by stage (cubature, solid angle, radius, map
        // integral, exposure, correction factor)
        auto cmap = Fermi::cubature_convolve_map_psf(Nh, Nw, src.sph_locs, psf.uPsf, obs.skygeom);
        Fermi::ModelMap::scale_map_by_solid_angle(cmap, obs.skygeom);
        const auto [radius, in_fov] =
            Fermi::ModelMap::psf_boundary_radius(Nh, Nw, src.sph_locs, obs.skygeom);
        const auto invmi =
            Fermi::ModelMap::map_integral(cmap, src.sph_locs, obs.skygeom, radius, in_fov);
        Fermi::ModelMap::scale_map_by_exposure(cmap, exp.exposure);
        const auto cf = Fermi::ModelMap::map_correction_factor(invmi, radius, in_fov, psf.uPsf,
                                                               psf.partial_psf_integral);
        Fermi::ModelMap::scale_map_by_correction_factors(cmap, cf);

        // SkyGeom on the pixel centres of the first row
        Fermi::Tensor<double, 2> sph(Nw, 2), pix(Nw, 2);
        for (long j = 0; j < Nw; ++j) {
            const auto s = obs.skygeom.pix2sph({1.0, (double)(j + 1)});
            sph[j, 0] = s[0];
            sph[j, 1] = s[1];
            const auto p = obs.skygeom.sph2pix(s);
            pix[j, 0] = p[0];
            pix[j, 1] = p[1];
        }
        Fermi::Tensor<double, 1> fov(in_fov.size());
        for (size_t s = 0; s < in_fov.size(); ++s) fov[s] = in_fov[s] ? 1.0 : 0.0;

        Arrays out;
        out["exposure"] = from_tensor(exp.exposure);
        out["ecb"] = from_tensor(exp.src_exposure_cosbins);
        out["upsf"] = from_tensor(psf.uPsf);
        out["parint"] = from_tensor(psf.partial_psf_integral);
        out["map"] = from_tensor(map);
        out["chain_upsf"] = from_tensor(mean);
        out["chain_parint"] = from_tensor(par);
        out["chain_map"] = from_tensor(cmap);
        out["radius"] = from_tensor(radius);
        out["in_fov"] = from_tensor(fov);
        out["inv_map_integral"] = from_tensor(invmi);
        out["correction"] = from_tensor(cf);
        out["row_sph"] = from_tensor(sph);
        out["row_pix"] = from_tensor(pix);
        out["aeff_front_params"] = from_tensor(irf.aeff_irf.front.effective_area.params);
        out["aeff_front_cosths"] = from_tensor(irf.aeff_irf.front.effective_area.cosths);
        out["aeff_front_logEs"] = from_tensor(irf.aeff_irf.front.effective_area.logEs);
        out["psf_back_params"] = from_tensor(irf.psf_irf.back.rpsf.params);
        out["ltf"] = from_tensor(irf.front_LTF);
        write_arrays(argv[a0 + 1], out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fermi_pipeline: %s\n", e.what());
        return 1;
    }
    return 0;
}
