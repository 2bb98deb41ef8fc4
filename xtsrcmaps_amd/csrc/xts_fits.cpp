// xts_fits.cpp -- the source-map FITS writer of the reference
// (fits/fits.write_src_model.cxx:10-66) without cfitsio.
//
// A FITS file is a sequence of HDUs; each is a header of 80-character cards
// padded with spaces to a multiple of 2880 bytes, then big-endian IEEE data
// padded with zeros to a multiple of 2880 bytes.  Every HDU carries the
// DATASUM and CHECKSUM cards of the FITS checksum convention: DATASUM is the
// 32-bit ones'-complement sum of the data as big-endian words, and CHECKSUM
// is the 16-character ASCII encoding of the complement of the header's sum
// plus DATASUM, which makes the sum of the whole HDU -0.
//
// The readers of the IRF tables (fits/fits.read_irf_pars.cxx:12-126) and of
// the livetime cube (fits/fits.read_expcube.cxx:8-116) parse binary-table
// HDUs the same way: header cards to keywords, TFORMn to column types and
// repeat counts, big-endian fields to host values.
#include <algorithm>
#include <cctype>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cerrno>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "xtsrcmaps_amd.h"

namespace xts {
int fail(int code, const char* fmt, ...);

namespace {

constexpr size_t kBlock = 2880;
constexpr size_t kCard = 80;

// ------------------------------------------------------------------- cards
std::string pad_card(std::string c) {
    c.resize(kCard, ' ');
    return c;
}

std::string key8(const char* key) {
    std::string k(key);
    k.resize(8, ' ');
    return k;
}

// Fixed format: value right-justified in columns 11-30.
std::string card_value(const char* key, const std::string& value, const char* comment) {
    std::string c = key8(key) + "= ";
    char v[32];
    std::snprintf(v, sizeof v, "%20s", value.c_str());
    c += v;
    if (comment && *comment) c += std::string(" / ") + comment;
    return pad_card(c);
}

std::string card_logical(const char* key, bool v, const char* comment = nullptr) {
    return card_value(key, v ? "T" : "F", comment);
}

std::string card_int(const char* key, long long v, const char* comment = nullptr) {
    return card_value(key, std::to_string(v), comment);
}

std::string card_real(const char* key, double v, const char* comment = nullptr) {
    // the shortest of 15 to 17 significant digits that reads back as v
    char b[40];
    for (int digits = 15; digits <= 17; ++digits) {
        std::snprintf(b, sizeof b, "%.*G", digits, v);
        if (std::strtod(b, nullptr) == v) break;
    }
    std::string s(b);
    if (s.find_first_of(".EN") == std::string::npos) s += ".";   // a real, not an integer
    return card_value(key, s, comment);
}

// String value: a quote in column 11, the text (quotes doubled) padded to
// at least 8 characters, a closing quote.
std::string card_str(const char* key, const std::string& v, const char* comment = nullptr) {
    std::string q;
    for (char ch : v) {
        q += ch;
        if (ch == '\'') q += '\'';
    }
    if (q.size() < 8) q.resize(8, ' ');
    std::string c = key8(key) + "= '" + q + "'";
    if (c.size() < 30) c.resize(30, ' ');
    if (comment && *comment) c += std::string(" / ") + comment;
    return pad_card(c);
}

// ---------------------------------------------------------------- checksum
// Ones'-complement sum of big-endian 32-bit words (len a multiple of 4).
uint32_t ones_sum(const unsigned char* p, size_t len, uint32_t start = 0) {
    uint64_t s = start;
    for (size_t i = 0; i < len; i += 4) {
        s += ((uint64_t)p[i] << 24) | ((uint64_t)p[i + 1] << 16) | ((uint64_t)p[i + 2] << 8) |
             (uint64_t)p[i + 3];
        if (s >> 32) s = (s & 0xffffffffu) + (s >> 32);
    }
    while (s >> 32) s = (s & 0xffffffffu) + (s >> 32);
    return (uint32_t)s;
}

uint32_t ones_add(uint32_t a, uint32_t b) {
    uint64_t s = (uint64_t)a + b;
    while (s >> 32) s = (s & 0xffffffffu) + (s >> 32);
    return (uint32_t)s;
}

// 16 characters whose words, added to a sum, give -0: each byte of the
// complement splits into four characters of about byte / 4 + '0' (the
// remainder on the first), pairs are nudged off the punctuation between the
// digits and the letters, and the string is rotated right by one character
// so that the 4-byte words line up with the card's alignment.
void checksum_ascii(uint32_t sum, char out[16]) {
    static const int kExclude[] = {0x3a, 0x3b, 0x3c, 0x3d, 0x3e, 0x3f, 0x40,
                                   0x5b, 0x5c, 0x5d, 0x5e, 0x5f, 0x60};
    const uint32_t value = ~sum;
    char asc[16];
    for (int i = 0; i < 4; ++i) {
        const int byte = (int)((value >> (24 - 8 * i)) & 0xff);
        const int q = byte / 4 + 0x30, r = byte % 4;
        int ch[4] = {q + r, q, q, q};
        for (bool moved = true; moved;) {
            moved = false;
            for (int x : kExclude)
                for (int j = 0; j < 4; j += 2)
                    if (ch[j] == x || ch[j + 1] == x) {
                        ++ch[j];
                        --ch[j + 1];
                        moved = true;
                    }
        }
        for (int j = 0; j < 4; ++j) asc[4 * j + i] = (char)ch[j];
    }
    for (int i = 0; i < 16; ++i) out[i] = asc[(i + 15) % 16];
}

// --------------------------------------------------------------------- HDU
struct Hdu {
    std::vector<std::string> cards;   // without END, CHECKSUM, DATASUM
    std::vector<unsigned char> data;  // big-endian, unpadded
};

// Header and data blocks of one HDU with its DATASUM and CHECKSUM.
std::vector<unsigned char> encode_hdu(const Hdu& h) {
    std::vector<unsigned char> data = h.data;
    data.resize((data.size() + kBlock - 1) / kBlock * kBlock, 0);
    const uint32_t datasum = ones_sum(data.data(), data.size());
    std::string hdr;
    for (auto& c : h.cards) hdr += c;
    const size_t ck_pos = hdr.size();
    hdr += card_str("CHECKSUM", "0000000000000000", "HDU checksum");
    hdr += card_str("DATASUM", std::to_string(datasum), "data unit checksum");
    hdr += pad_card("END");
    hdr.resize((hdr.size() + kBlock - 1) / kBlock * kBlock, ' ');
    const uint32_t sum =
        ones_add(ones_sum(reinterpret_cast<const unsigned char*>(hdr.data()), hdr.size()),
                 datasum);
    char enc[16];
    checksum_ascii(sum, enc);
    std::memcpy(&hdr[ck_pos + 11], enc, 16);   // between the quotes of the value
    std::vector<unsigned char> out(hdr.begin(), hdr.end());
    out.insert(out.end(), data.begin(), data.end());
    return out;
}

void put_be(unsigned char* p, double v, bool f32) {
    if (f32) {
        const float f = (float)v;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        for (int k = 0; k < 4; ++k) p[k] = (unsigned char)(u >> (24 - 8 * k));
    } else {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        for (int k = 0; k < 8; ++k) p[k] = (unsigned char)(u >> (56 - 8 * k));
    }
}

void wcs_cards(std::vector<std::string>& c, const xts_wcs* w) {
    const char* proj = w->proj == XTS_PROJ_AIT ? "AIT" : "CAR";
    const std::string lon = w->is_galactic ? "GLON" : "RA--";
    const std::string lat = w->is_galactic ? "GLAT" : "DEC-";
    c.push_back(card_str("CTYPE1", lon + "-" + proj));
    c.push_back(card_real("CRPIX1", w->crpix1));
    c.push_back(card_real("CRVAL1", w->crval1));
    c.push_back(card_real("CDELT1", w->cdelt1));
    c.push_back(card_str("CUNIT1", "deg"));
    c.push_back(card_str("CTYPE2", lat + "-" + proj));
    c.push_back(card_real("CRPIX2", w->crpix2));
    c.push_back(card_real("CRVAL2", w->crval2));
    c.push_back(card_real("CDELT2", w->cdelt2));
    c.push_back(card_str("CUNIT2", "deg"));
    c.push_back(card_real("CROTA2", w->crota2));
    if (!w->is_galactic) c.push_back(card_real("EQUINOX", 2000.0));
}

// ------------------------------------------------------------------ reading
struct Column {
    char type;        // TFORM letter
    int64_t repeat;   // elements per row
    int64_t offset;   // byte offset in the row
};

struct TableHdu {
    std::map<std::string, std::string> keys;   // card values, strings unquoted
    std::vector<Column> cols;
    int64_t row_bytes = 0, nrows = 0;
    const unsigned char* data = nullptr;
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

// Value field of a card (columns 11-80): a quoted string (quotes doubled
// inside, trailing blanks dropped) or the text before a '/' comment.
std::string card_field(const std::string& v) {
    std::string t = trim(v);
    if (!t.empty() && t[0] == '\'') {
        std::string out;
        for (size_t i = 1; i < t.size(); ++i) {
            if (t[i] == '\'') {
                if (i + 1 < t.size() && t[i + 1] == '\'') {
                    out += '\'';
                    ++i;
                    continue;
                }
                break;
            }
            out += t[i];
        }
        return trim(out);
    }
    return trim(t.substr(0, t.find('/')));
}

int64_t type_bytes(char t) {
    switch (t) {
        case 'L': case 'B': case 'A': return 1;
        case 'I': return 2;
        case 'J': case 'E': return 4;
        case 'K': case 'D': case 'C': return 8;
        case 'M': return 16;
        default: return -1;
    }
}

// The primary HDU: its keywords and data.
struct PrimaryHdu {
    std::map<std::string, std::string> keys;
    const unsigned char* data = nullptr;
    int64_t bytes = 0;
};

// The binary tables of a FITS file held in memory, by EXTNAME (upper case),
// and (prim) its primary HDU.
int scan_tables(const std::vector<unsigned char>& buf, std::map<std::string, TableHdu>* out,
                PrimaryHdu* prim = nullptr) {
    size_t pos = 0;
    bool first = true;
    while (pos + kBlock <= buf.size()) {
        std::map<std::string, std::string> keys;
        bool end = false;
        while (!end) {
            if (pos + kBlock > buf.size()) return fail(XTS_ERR_INVALID, "FITS header truncated");
            for (size_t i = 0; i < kBlock && !end; i += kCard) {
                const std::string card(reinterpret_cast<const char*>(&buf[pos + i]), kCard);
                const std::string key = trim(card.substr(0, 8));
                if (key == "END") end = true;
                else if (card.compare(8, 2, "= ") == 0) keys[key] = card_field(card.substr(10));
            }
            pos += kBlock;
        }
        if (first && keys.count("SIMPLE") == 0) return fail(XTS_ERR_INVALID, "not a FITS file");
        first = false;
        // integer keywords: digits only (optional sign), in range
        bool bad_num = false;
        auto num = [&](const char* k, int64_t dflt) -> int64_t {
            auto it = keys.find(k);
            if (it == keys.end()) return dflt;
            const char* c = it->second.c_str();
            char* e = nullptr;
            errno = 0;
            const long long v = std::strtoll(c, &e, 10);
            if (e == c || *e != '\0' || errno == ERANGE) bad_num = true;
            return v;
        };
        const int64_t naxis = num("NAXIS", 0), bitpix = num("BITPIX", 8);
        const int64_t pcount = num("PCOUNT", 0), gcount = num("GCOUNT", 1);
        if (bad_num) return fail(XTS_ERR_INVALID, "FITS header: malformed integer keyword");
        if (naxis < 0 || naxis > 999)
            return fail(XTS_ERR_INVALID, "FITS header: NAXIS = %lld", (long long)naxis);
        if (bitpix != 8 && bitpix != 16 && bitpix != 32 && bitpix != 64 && bitpix != -32 &&
            bitpix != -64)
            return fail(XTS_ERR_INVALID, "FITS header: BITPIX = %lld", (long long)bitpix);
        if (pcount < 0 || gcount < 0)
            return fail(XTS_ERR_INVALID, "FITS header: PCOUNT = %lld, GCOUNT = %lld",
                        (long long)pcount, (long long)gcount);
        // data bytes = |BITPIX|/8 * prod(NAXISn), + PCOUNT, x GCOUNT, checked
        // for overflow and against what the file holds after this header
        int64_t n = naxis > 0 ? (bitpix < 0 ? -bitpix : bitpix) / 8 : 0;
        for (int64_t k = 1; k <= naxis; ++k) {
            const int64_t ax = num(("NAXIS" + std::to_string(k)).c_str(), 0);
            if (bad_num || ax < 0)
                return fail(XTS_ERR_INVALID, "FITS header: NAXIS%lld = %lld", (long long)k,
                            (long long)ax);
            if (__builtin_mul_overflow(n, ax, &n))
                return fail(XTS_ERR_INVALID, "FITS header: data size overflows");
        }
        if (__builtin_add_overflow(n, pcount, &n) ||
            __builtin_mul_overflow(n, std::max<int64_t>(gcount, 1), &n))
            return fail(XTS_ERR_INVALID, "FITS header: data size overflows");
        if ((uint64_t)n > buf.size() - pos) return fail(XTS_ERR_INVALID, "FITS data truncated");
        if (prim && !keys.count("XTENSION")) {
            prim->keys = keys;
            prim->data = buf.data() + pos;
            prim->bytes = n;
        }
        if (keys.count("XTENSION") && upper(keys["XTENSION"]) == "BINTABLE") {
            TableHdu t;
            t.row_bytes = num("NAXIS1", 0);
            t.nrows = num("NAXIS2", 0);
            const int64_t tfields = num("TFIELDS", 0);
            if (naxis != 2 || t.row_bytes < 0 || t.nrows < 0 || tfields < 0 || tfields > 999)
                return fail(XTS_ERR_INVALID, "binary table: NAXIS = %lld, TFIELDS = %lld",
                            (long long)naxis, (long long)tfields);
            // the rows lie inside the HDU's data (the heap follows them)
            int64_t table_bytes;
            if (__builtin_mul_overflow(t.row_bytes, t.nrows, &table_bytes) || table_bytes > n)
                return fail(XTS_ERR_INVALID, "binary table: rows exceed the data");
            t.data = buf.data() + pos;
            int64_t off = 0;
            for (int64_t c = 1; c <= tfields; ++c) {
                const std::string f = keys["TFORM" + std::to_string(c)];
                size_t i = 0;
                while (i < f.size() && std::isdigit((unsigned char)f[i])) ++i;
                if (i > 18) return fail(XTS_ERR_INVALID, "column format '%s' too long", f.c_str());
                Column col{i < f.size() ? (char)std::toupper((unsigned char)f[i]) : '?',
                           i > 0 ? std::strtoll(f.substr(0, i).c_str(), nullptr, 10) : 1, off};
                const int64_t w = type_bytes(col.type);
                if (w < 0) return fail(XTS_ERR_INVALID, "unsupported column format '%s'", f.c_str());
                int64_t cb;
                if (__builtin_mul_overflow(w, col.repeat, &cb) || __builtin_add_overflow(off, cb, &off))
                    return fail(XTS_ERR_INVALID, "column format '%s' overflows", f.c_str());
                t.cols.push_back(col);
            }
            if (off != t.row_bytes)
                return fail(XTS_ERR_INVALID, "row of %lld bytes, columns of %lld",
                            (long long)t.row_bytes, (long long)off);
            t.keys = std::move(keys);
            const std::string name = upper(t.keys.count("EXTNAME") ? t.keys["EXTNAME"] : "");
            if (!out->count(name)) (*out)[name] = std::move(t);
        }
        pos += ((size_t)n + kBlock - 1) / kBlock * kBlock;
    }
    return XTS_OK;
}

int load_file(const char* path, std::vector<unsigned char>* buf) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(XTS_ERR_INVALID, "cannot open '%s'", path);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return fail(XTS_ERR_INVALID, "cannot size '%s'", path);
    }
    buf->resize(n > 0 ? (size_t)n : 0);
    const size_t got = buf->empty() ? 0 : std::fread(buf->data(), 1, buf->size(), f);
    std::fclose(f);
    if (got != buf->size()) return fail(XTS_ERR_INVALID, "cannot read '%s'", path);
    return XTS_OK;
}

int find_table(const char* path, const char* name, std::vector<unsigned char>* buf,
               TableHdu* t) {
    if (load_file(path, buf) != XTS_OK) return XTS_ERR_INVALID;
    std::map<std::string, TableHdu> tables;
    if (scan_tables(*buf, &tables) != XTS_OK) return XTS_ERR_INVALID;
    auto it = tables.find(upper(trim(name)));
    if (it == tables.end())
        return fail(XTS_ERR_INVALID, "no binary table '%s' in '%s'", name, path);
    *t = it->second;
    return XTS_OK;
}

uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

// Element k of a column in a row, as float (from E or D fields).
float field_float(const TableHdu& t, const Column& c, int64_t row, int64_t k) {
    const unsigned char* p = t.data + row * t.row_bytes + c.offset;
    if (c.type == 'E') {
        const uint32_t u = be32(p + 4 * k);
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
    const uint64_t u = ((uint64_t)be32(p + 8 * k) << 32) | be32(p + 8 * k + 4);
    double d;
    std::memcpy(&d, &u, 8);
    return (float)d;
}

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};

}  // namespace
}  // namespace xts

using namespace xts;

extern "C" int xts_write_src_model(const char* path, const double* map, int64_t S, int64_t H,
                                   int64_t W, int64_t E, const char* const* names, int32_t flags,
                                   const xts_wcs* wcs) {
    if (!path || !map || !names) return fail(XTS_ERR_INVALID, "xts_write_src_model: null pointer");
    if (S <= 0 || H <= 0 || W <= 0 || E <= 0)
        return fail(XTS_ERR_INVALID, "xts_write_src_model: bad sizes S=%lld H=%lld W=%lld E=%lld",
                    (long long)S, (long long)H, (long long)W, (long long)E);
    if ((flags & ~(XTS_LAYOUT_SEHW | XTS_OUT_F32)) != 0)
        return fail(XTS_ERR_INVALID, "xts_write_src_model: unknown flags 0x%x", flags);
    if (wcs && (wcs->naxis1 != H || wcs->naxis2 != W))
        return fail(XTS_ERR_INVALID, "xts_write_src_model: wcs is %lldx%lld, map %lldx%lld",
                    (long long)wcs->naxis1, (long long)wcs->naxis2, (long long)H, (long long)W);
    for (int64_t s = 0; s < S; ++s)
        if (!names[s]) return fail(XTS_ERR_INVALID, "xts_write_src_model: null name %lld", (long long)s);
    const bool sehw = (flags & XTS_LAYOUT_SEHW) != 0, f32 = (flags & XTS_OUT_F32) != 0;
    const size_t esz = f32 ? 4 : 8;
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "wb"));
    if (!f) return fail(XTS_ERR_INVALID, "xts_write_src_model: cannot open '%s'", path);
    auto put = [&](const std::vector<unsigned char>& b) {
        return std::fwrite(b.data(), 1, b.size(), f.get()) == b.size();
    };
    // a primary HDU without data, named as the reference names it
    Hdu prim;
    prim.cards = {card_logical("SIMPLE", true, "conforms to FITS standard"),
                  card_int("BITPIX", f32 ? -32 : -64), card_int("NAXIS", 0),
                  card_logical("EXTEND", true), card_str("HDUNAME", "PRIMARY")};
    if (!put(encode_hdu(prim))) return fail(XTS_ERR_INVALID, "xts_write_src_model: write failed");
    const size_t plane = (size_t)H * W;
    for (int64_t s = 0; s < S; ++s) {
        Hdu img;
        img.cards = {card_str("XTENSION", "IMAGE", "image extension"),
                     card_int("BITPIX", f32 ? -32 : -64), card_int("NAXIS", 3),
                     card_int("NAXIS1", H), card_int("NAXIS2", W), card_int("NAXIS3", E),
                     card_int("PCOUNT", 0), card_int("GCOUNT", 1),
                     card_str("EXTNAME", names[s], "source name")};
        if (wcs) wcs_cards(img.cards, wcs);
        // the source's cube as [E][W][H], H fastest (FITS axis 1)
        img.data.resize((size_t)E * plane * esz);
        const double* m = map + (size_t)s * plane * E;
        unsigned char* p = img.data.data();
        for (int64_t e = 0; e < E; ++e)
            for (int64_t w = 0; w < W; ++w)
                for (int64_t h = 0; h < H; ++h, p += esz) {
                    const double v = sehw ? m[((size_t)e * H + h) * W + w]
                                          : m[((size_t)h * W + w) * E + e];
                    put_be(p, v, f32);
                }
        if (!put(encode_hdu(img)))
            return fail(XTS_ERR_INVALID, "xts_write_src_model: write failed");
    }
    if (std::fflush(f.get()) != 0) return fail(XTS_ERR_INVALID, "xts_write_src_model: write failed");
    return XTS_OK;
}

// fits::read_allsky_cropped's input (fits/fits.read_allsky_cropped.cpp:13-41,
// 150-216): the primary image of a MapCube file, its WCS (readWcsParams's
// keywords; CROTA2 0 when absent, the frame from CTYPE1) and the ENERGIES
// table's first column.
extern "C" int xts_read_mapcube(const char* path, xts_wcs* wcs, int64_t* planes, float* data,
                                double* energies, int64_t* n_energies) {
    if (!path || !wcs || !planes) return fail(XTS_ERR_INVALID, "xts_read_mapcube: null pointer");
    std::vector<unsigned char> buf;
    if (load_file(path, &buf) != XTS_OK) return XTS_ERR_INVALID;
    std::map<std::string, TableHdu> tables;
    PrimaryHdu prim;
    if (scan_tables(buf, &tables, &prim) != XTS_OK) return XTS_ERR_INVALID;
    auto& k = prim.keys;
    auto inum = [&](const char* key, int64_t dflt) -> int64_t {
        return k.count(key) ? std::strtoll(k[key].c_str(), nullptr, 10) : dflt;
    };
    bool bad = false;
    auto real = [&](const char* key, bool required, double dflt) -> double {
        if (!k.count(key)) {
            if (required) bad = true;
            return dflt;
        }
        const char* c = k[key].c_str();
        char* e = nullptr;
        const double v = std::strtod(c, &e);
        if (e == c || *e != '\0' || !std::isfinite(v)) bad = true;
        return v;
    };
    const int64_t naxis = inum("NAXIS", 0), bitpix = inum("BITPIX", 8);
    const int64_t n1 = inum("NAXIS1", 0), n2 = inum("NAXIS2", 0);
    const int64_t n3 = naxis >= 3 ? inum("NAXIS3", 1) : 1;
    if ((naxis != 2 && naxis != 3) || n1 <= 0 || n2 <= 0 || n3 <= 0 || n1 > 65536 ||
        n2 > 65536 || n3 > 65536)
        return fail(XTS_ERR_INVALID, "'%s': primary image of NAXIS %lld (%lld x %lld x %lld)",
                    path, (long long)naxis, (long long)n1, (long long)n2, (long long)n3);
    if (bitpix != -32 && bitpix != -64)
        return fail(XTS_ERR_INVALID, "'%s': BITPIX %lld (real images only)", path,
                    (long long)bitpix);
    const std::string ctype = upper(k.count("CTYPE1") ? k["CTYPE1"] : "");
    const std::string proj = ctype.size() >= 8 ? ctype.substr(5, 3) : "";
    if (proj != "CAR" && proj != "AIT")
        return fail(XTS_ERR_INVALID, "'%s': CTYPE1 '%s' (CAR or AIT only)", path, ctype.c_str());
    xts_wcs w{};
    w.naxis1 = n1;
    w.naxis2 = n2;
    w.crpix1 = real("CRPIX1", true, 0.0);
    w.crpix2 = real("CRPIX2", true, 0.0);
    w.crval1 = real("CRVAL1", true, 0.0);
    w.crval2 = real("CRVAL2", true, 0.0);
    w.cdelt1 = real("CDELT1", true, 0.0);
    w.cdelt2 = real("CDELT2", true, 0.0);
    w.crota2 = real("CROTA2", false, 0.0);
    w.proj = proj == "CAR" ? XTS_PROJ_CAR : XTS_PROJ_AIT;
    w.is_galactic = ctype.compare(0, 4, "GLON") == 0 ? 1 : 0;
    if (bad || w.cdelt1 == 0.0 || w.cdelt2 == 0.0)
        return fail(XTS_ERR_INVALID, "'%s': missing or malformed CRPIX/CRVAL/CDELT", path);
    *wcs = w;
    *planes = n3;
    if (data) {
        const unsigned char* p = prim.data;
        const size_t n = (size_t)n1 * n2 * n3;
        for (size_t i = 0; i < n; ++i) {
            if (bitpix == -32) {
                const uint32_t u = be32(p + 4 * i);
                std::memcpy(&data[i], &u, 4);
            } else {
                const uint64_t u = ((uint64_t)be32(p + 8 * i) << 32) | be32(p + 8 * i + 4);
                double d;
                std::memcpy(&d, &u, 8);
                data[i] = (float)d;
            }
        }
    }
    if (n_energies || energies) {
        auto it = tables.find("ENERGIES");
        int64_t ne = 0;
        if (it != tables.end() && !it->second.cols.empty()) {
            const TableHdu& t = it->second;
            const Column& c = t.cols[0];
            if (c.type != 'E' && c.type != 'D')
                return fail(XTS_ERR_INVALID, "'%s': ENERGIES column 1 is not real", path);
            ne = t.nrows * c.repeat;
            if (energies)
                for (int64_t i = 0; i < ne; ++i) {
                    const unsigned char* q = t.data + (i / c.repeat) * t.row_bytes + c.offset;
                    const int64_t j = i % c.repeat;
                    if (c.type == 'E') {
                        const uint32_t u = be32(q + 4 * j);
                        float f;
                        std::memcpy(&f, &u, 4);
                        energies[i] = f;
                    } else {
                        const uint64_t u = ((uint64_t)be32(q + 8 * j) << 32) | be32(q + 8 * j + 4);
                        std::memcpy(&energies[i], &u, 8);
                    }
                }
        }
        if (n_energies) *n_energies = ne;
    }
    return XTS_OK;
}

extern "C" int xts_read_irf_pars(const char* path, const char* tblname, int64_t* ncols,
                                 int64_t* extents, int64_t cap_cols, int64_t* nrows,
                                 int64_t* row_width, float* rowdata) {
    if (!path || !tblname || !ncols || !nrows || !row_width)
        return fail(XTS_ERR_INVALID, "xts_read_irf_pars: null pointer");
    std::vector<unsigned char> buf;
    TableHdu t;
    if (find_table(path, tblname, &buf, &t) != XTS_OK) return XTS_ERR_INVALID;
    int64_t width = 0;
    for (size_t c = 0; c < t.cols.size(); ++c) {
        if (t.cols[c].type != 'E')
            return fail(XTS_ERR_INVALID, "table '%s' column %zu is not float32 ('%c')", tblname,
                        c + 1, t.cols[c].type);
        if (extents && (int64_t)c < cap_cols) extents[c] = t.cols[c].repeat;
        width += t.cols[c].repeat;
    }
    *ncols = (int64_t)t.cols.size();
    *nrows = t.nrows;
    *row_width = width;
    if (rowdata) {
        // each row: the columns' elements in column order
        for (int64_t r = 0; r < t.nrows; ++r) {
            float* o = rowdata + r * width;
            for (const Column& c : t.cols)
                for (int64_t k = 0; k < c.repeat; ++k) *o++ = field_float(t, c, r, k);
        }
    }
    return XTS_OK;
}

extern "C" int xts_read_expcube(const char* path, const char* tblname, xts_expcube_info* info,
                                float* cosbins, float* ra, float* dec) {
    if (!path || !tblname || !info) return fail(XTS_ERR_INVALID, "xts_read_expcube: null pointer");
    std::vector<unsigned char> buf;
    TableHdu t;
    if (find_table(path, tblname, &buf, &t) != XTS_OK) return XTS_ERR_INVALID;
    for (const char* k : {"NSIDE", "NBRBINS", "COSMIN", "ORDERING", "COORDSYS", "THETABIN"})
        if (!t.keys.count(k))
            return fail(XTS_ERR_INVALID, "xts_read_expcube: no %s keyword in '%s'", k, tblname);
    const std::string ordering = t.keys["ORDERING"], coordsys = t.keys["COORDSYS"];
    if (ordering != "NESTED")
        return fail(XTS_ERR_INVALID, "livetime cube ordering %s is unsupported (NESTED only)",
                    ordering.c_str());
    if (coordsys != "EQU")
        return fail(XTS_ERR_INVALID, "livetime cube COORDSYS %s is unsupported (EQU only)",
                    coordsys.c_str());
    const int64_t nside = std::strtoll(t.keys["NSIDE"].c_str(), nullptr, 10);
    const int64_t nbr = std::strtoll(t.keys["NBRBINS"].c_str(), nullptr, 10);
    // bounds keep every product below 2^62 (row counts and repeats are
    // bounded by the file's size, checked when the table was scanned)
    if (nside <= 0 || nside > (1 << 13) || nbr <= 0 || nbr > (1 << 16) || t.cols.size() < 3 ||
        t.cols[0].repeat <= 0 || t.cols[1].repeat <= 0 || t.cols[2].repeat <= 0)
        return fail(XTS_ERR_INVALID, "livetime cube '%s': NSIDE %lld, NBRBINS %lld or its "
                    "columns out of range", tblname, (long long)nside, (long long)nbr);
    const int64_t npix = 12 * nside * nside;
    if (t.cols[0].repeat * t.nrows < nbr * npix || t.cols[1].repeat * t.nrows < npix ||
        t.cols[2].repeat * t.nrows < npix)
        return fail(XTS_ERR_INVALID, "livetime cube '%s' does not hold 12 nside^2 x %lld bins",
                    tblname, (long long)nbr);
    for (int c = 0; c < 3; ++c)
        if (t.cols[c].type != 'E' && t.cols[c].type != 'D')
            return fail(XTS_ERR_INVALID, "livetime cube column %d is not real", c + 1);
    info->nside = nside;
    info->nbrbins = nbr;
    info->cosmin = std::strtod(t.keys["COSMIN"].c_str(), nullptr);
    info->thetabin = t.keys["THETABIN"] != "COSTHETA" ? 1 : 0;
    std::snprintf(info->ordering, sizeof info->ordering, "%s", ordering.c_str());
    std::snprintf(info->coordsys, sizeof info->coordsys, "%s", coordsys.c_str());
    // column n read as one run of elements across rows (row-major), as
    // fits_read_col does
    auto column = [&](int c, int64_t n, float* out) {
        const Column& col = t.cols[c];
        for (int64_t i = 0; i < n; ++i) out[i] = field_float(t, col, i / col.repeat, i % col.repeat);
    };
    if (cosbins) column(0, nbr * npix, cosbins);
    if (ra) column(1, npix, ra);
    if (dec) column(2, npix, dec);
    return XTS_OK;
}
