// The diffuse-source steps of the reference's C++ API, as this library
// implements them: Fermi::ModelMap::diffuse_src_model_map_wcs,
// Fermi::Psf::make_kernel and Fermi::FFT::convolve_map_psf, called the way
// model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97 calls them.
//
//   diffuse_driver <mapcube.fits> <out.bin>
//
// Writes the inputs it made and the three results to out.bin as a sequence of
// arrays: int64 rank, int64 extents, float64 data (row-major).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "xtsrcmaps/model_map/model_map.hxx"
#include "xtsrcmaps/psf/psf.hxx"
#include "xtsrcmaps/sky_geom/sky_geom.hxx"
#include "xtsrcmaps/tensor/tensor.hpp"

namespace {

template <std::size_t R>
void put(std::FILE* f, Fermi::Tensor<double, R> const& t) {
    const int64_t rank = (int64_t)R;
    std::fwrite(&rank, 8, 1, f);
    size_t n = 1;
    for (std::size_t k = 0; k < R; ++k) {
        const int64_t x = (int64_t)t.extent(k);
        std::fwrite(&x, 8, 1, f);
        n *= t.extent(k);
    }
    std::fwrite(t.data(), 8, n, f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: diffuse_driver <mapcube.fits> <out.bin>\n");
        return 2;
    }
    try {
        const size_t Nh = 40, Nw = 36, E = 4;
        // a celestial CAR image at 0.25 deg around (193.98, -5.82)
        Fermi::SkyGeom<double> roi(40, 36, "CAR", false, 20.5, 18.5, 193.98, -5.82, -0.25, 0.25,
                                   0.0);
        Fermi::Tensor<double, 2> src_sph(1, 2);
        src_sph[0, 0] = 193.98;
        src_sph[0, 1] = -5.82;
        // the diffuse source's PSF table and exposure (row 0)
        const auto sep = Fermi::Psf::separations();
        Fermi::Tensor<double, 3> uPsf(1, Fermi::Psf::sep_arr_len, E);
        Fermi::Tensor<double, 3> parint(1, Fermi::Psf::sep_arr_len, E);
        for (size_t d = 0; d < Fermi::Psf::sep_arr_len; ++d)
            for (size_t e = 0; e < E; ++e) {
                const double x = sep[d] / (0.5 * (e + 1.0));
                uPsf[0, d, e] = std::pow(1.0 + x * x / 5.0, -2.5);
            }
        Fermi::Tensor<double, 2> expo(1, E);
        for (size_t e = 0; e < E; ++e) expo[0, e] = 1e10 * (1.0 + e);
        const std::vector<std::string> names = {argv[1]};
        auto map = Fermi::ModelMap::diffuse_src_model_map_wcs(Nh, Nw, roi, src_sph, names, expo,
                                                              uPsf, parint, 0.5);
        auto kernel = Fermi::Psf::make_kernel(15, roi, uPsf);
        // FFT::convolve_map_psf on a deterministic image and kernel
        Fermi::Tensor<double, 3> img(2, 23, 17), ker(2, 9, 5);
        uint64_t st = 12345;
        auto next = [&st]() {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            return (double)(st >> 11) * (1.0 / 9007199254740992.0);
        };
        for (size_t c = 0; c < 2; ++c) {
            for (size_t h = 0; h < 23; ++h)
                for (size_t w = 0; w < 17; ++w) img[c, h, w] = next() - 0.5;
            for (size_t h = 0; h < 9; ++h)
                for (size_t w = 0; w < 5; ++w) ker[c, h, w] = next();
        }
        auto conv = Fermi::FFT::convolve_map_psf(img, ker);
        std::FILE* f = std::fopen(argv[2], "wb");
        if (!f) {
            std::fprintf(stderr, "cannot open %s\n", argv[2]);
            return 1;
        }
        put(f, uPsf);
        put(f, expo);
        put(f, map);
        put(f, kernel);
        put(f, img);
        put(f, ker);
        put(f, conv);
        std::fclose(f);
    } catch (std::exception const& e) {
        std::fprintf(stderr, "diffuse_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
