// fits_check.cpp -- host driver of the FITS entry points (xts_fits.cpp) for
// a sanitizer build: AddressSanitizer and UndefinedBehaviorSanitizer on the
// hand-written FITS parser and writer, with no device code linked.
//
//   fits_check irf <file> <table>     xts_read_irf_pars (sizes, then rows)
//   fits_check expcube <file> <table> xts_read_expcube (keywords, then arrays)
//   fits_check write <file>           xts_write_src_model of a small map,
//                                     then its HDU read back as a table
//   fits_check mapcube <file>         xts_read_mapcube (sizes, then the
//                                     planes and the energies)
// Prints "rc <code>" per call; exit status 0 unless the sanitizers abort.
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xtsrcmaps_amd.h"

namespace xts {
// the library's error reporter lives in the device part; a host stand-in here
int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stdout, fmt, ap);
    va_end(ap);
    std::fputc('\n', stdout);
    return code;
}
}  // namespace xts

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    if (mode == "irf" && argc == 4) {
        int64_t ncols = 0, nrows = 0, width = 0, ext[64] = {};
        int rc = xts_read_irf_pars(argv[2], argv[3], &ncols, ext, 64, &nrows, &width, nullptr);
        std::printf("rc %d\n", rc);
        if (rc == 0) {
            std::vector<float> rows((size_t)(nrows * width) + 1);
            rc = xts_read_irf_pars(argv[2], argv[3], &ncols, ext, 64, &nrows, &width, rows.data());
            double sum = 0;
            for (float v : rows) sum += v;
            std::printf("rc %d rows %lld width %lld sum %g\n", rc, (long long)nrows,
                        (long long)width, sum);
        }
        return 0;
    }
    if (mode == "expcube" && argc == 4) {
        xts_expcube_info info{};
        int rc = xts_read_expcube(argv[2], argv[3], &info, nullptr, nullptr, nullptr);
        std::printf("rc %d\n", rc);
        if (rc == 0) {
            const size_t npix = (size_t)(12 * info.nside * info.nside);
            std::vector<float> cb(npix * (size_t)info.nbrbins), ra(npix), dec(npix);
            rc = xts_read_expcube(argv[2], argv[3], &info, cb.data(), ra.data(), dec.data());
            std::printf("rc %d nside %lld\n", rc, (long long)info.nside);
        }
        return 0;
    }
    if (mode == "mapcube") {
        xts_wcs w{};
        int64_t planes = 0, ne = 0;
        int rc = xts_read_mapcube(argv[2], &w, &planes, nullptr, nullptr, &ne);
        std::printf("rc %d\n", rc);
        if (rc == 0) {
            std::vector<float> data((size_t)(planes * w.naxis1 * w.naxis2));
            std::vector<double> en((size_t)ne);
            rc = xts_read_mapcube(argv[2], &w, &planes, data.data(), en.data(), &ne);
            double sum = 0, esum = 0;
            for (float v : data) sum += v;
            for (double v : en) esum += v;
            std::printf("rc %d planes %lld size %lld x %lld energies %lld sum %g esum %g\n", rc,
                        (long long)planes, (long long)w.naxis1, (long long)w.naxis2,
                        (long long)ne, sum, esum);
        }
        return 0;
    }
    if (mode == "write") {
        const int64_t S = 2, H = 5, W = 3, E = 4;
        std::vector<double> map((size_t)(S * H * W * E));
        for (size_t i = 0; i < map.size(); ++i) map[i] = 0.5 + (double)i;
        const char* names[] = {"SRC_A", "SRC'B"};
        xts_wcs w{H, W, 3.0, 2.0, 180.0, 0.0, -0.1, 0.1, 0.0, XTS_PROJ_CAR, 0};
        int rc = xts_write_src_model(argv[2], map.data(), S, H, W, E, names, XTS_OUT_F32, &w);
        std::printf("rc %d\n", rc);
        int64_t ncols, nrows, width;
        rc = xts_read_irf_pars(argv[2], "SRC_A", &ncols, nullptr, 0, &nrows, &width, nullptr);
        std::printf("rc %d (an image HDU is not a table)\n", rc);
        return 0;
    }
    return 2;
}
